// Host side of the GPU FFA search: builds the device execution layout of an
// FfaSearchPlan, owns workspace and the compact peak buffer, runs series
// tiles through gaffa_hip_ffa_search_batch, grows the peak buffer and repeats
// a tile on overflow (ffa_cuda_executor.cuh:385-531 in gaffa), and expands
// compact peaks with make_ffa_peak.
#include "gaffa/ffa_gpu.h"

#include <algorithm>
#include <cmath>
#include <parallel/algorithm>
#include <cstring>
#include <map>
#include <stdexcept>

namespace gaffa {
namespace {

constexpr std::size_t kAlign = 64;               // floats
// Most dynamic LDS a level launch is given: two workgroups per CU at least.
constexpr std::uint64_t kLevelLdsCap = 64 * 1024;
constexpr std::size_t kSubtreeMaxRows = 512;
constexpr std::size_t kSubtreeLdsBytes = 160 * 1024;

// LDS of a subtree launch besides its two row buffers (include/gaffa_hip.h,
// subtree_lds_bytes): the -0.0 zero rows and 1 KiB of slack.
std::size_t subtree_fixed_lds(std::size_t max_bins) {
  const std::size_t zero = (std::max<std::size_t>(256, max_bins + 64) + 63) / 64 * 64;
  return 4 * zero + 1024;
}

std::size_t plan_max_bins(const FfaSearchPlan& plan) {
  std::size_t b = 0;
  for (const FfaSearchTask& t : plan.tasks) b = std::max(b, t.bins);
  return b;
}

std::size_t tree_levels(std::size_t rows) {
  std::size_t levels = 0;
  for (std::size_t sz = rows; sz > 1; sz = sz - sz / 2) ++levels;
  return levels;
}

std::size_t align_up(std::size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

void validate_task(const FfaSearchTask& t, std::size_t n) {
  check_arg(t.input_nsamples == n, "FFA executor task input_nsamples must match time series size");
  check_arg(std::isfinite(t.downsample_factor) && t.downsample_factor >= 1.0,
            "FFA executor task downsample_factor must be finite and >= 1");
  check_arg(t.effective_tsamp > 0.0 && std::isfinite(t.effective_tsamp),
            "FFA executor task effective_tsamp must be finite and > 0");
  check_arg(t.bins > 1, "FFA executor task bins must be > 1");
  check_arg(t.rows != 0, "FFA executor task rows must be > 0");
  check_arg(t.rows_eval != 0 && t.rows_eval <= t.rows,
            "FFA executor task rows_eval must satisfy 0 < rows_eval <= rows");
  check_arg(t.prepared_nsamples != 0, "FFA executor task prepared_nsamples must be > 0");
  if (t.rows > std::numeric_limits<std::size_t>::max() / t.bins) {
    throw std::overflow_error("FFA executor shape size overflow");
  }
  check_arg(t.rows * t.bins <= t.prepared_nsamples,
            "FFA executor task rows * bins must be <= prepared_nsamples");
  if (t.downsample_factor == 1.0) {
    check_arg(t.prepared_nsamples == n,
              "FFA executor no-downsample task prepared_nsamples must match input size");
  } else {
    check_arg(t.prepared_nsamples == downsampled_size(n, t.downsample_factor),
              "FFA executor task prepared_nsamples must match downsampled_size");
  }
  check_arg(t.bins < 65536, "GPU FFA supports bins < 65536");
  check_arg(t.rows < (1ULL << 31), "GPU FFA supports rows < 2^31");
  // the kernels address a task's rows with 32-bit byte offsets
  check_arg(t.rows * t.bins < (1ULL << 30), "GPU FFA supports rows * bins < 2^30 per task");
}

void collect_nodes(std::uint32_t start, std::uint32_t size, std::uint32_t depth,
                   std::uint32_t target, std::uint32_t task, std::vector<gaffa_hip_ffa_item>& out) {
  if (depth == target) {
    out.push_back(gaffa_hip_ffa_item{task, start, size, 0});
    return;
  }
  if (size < 2) throw std::logic_error("FFA subtree depth below a leaf");
  const std::uint32_t h = size / 2;
  collect_nodes(start, h, depth + 1, target, task, out);
  collect_nodes(start + h, size - h, depth + 1, target, task, out);
}

}  // namespace

namespace {

std::uint32_t pick_rows(std::uint32_t want, std::uint32_t dflt, std::uint32_t max,
                        const char* message) {
  if (want == 0) return dflt;
  check_arg(want <= max, message);
  return want;
}

}  // namespace

GpuFfaExecutionPlan::GpuFfaExecutionPlan(const FfaSearchPlan& plan, FfaItemShape shape)
    : plan_(plan) {
  level_rows_ = pick_rows(shape.level_rows, gaffa_hip_ffa_level_rows(),
                          gaffa_hip_ffa_level_rows_max(), "FFA level rows per item too large");
  terminal_rows_ = pick_rows(shape.terminal_rows, gaffa_hip_ffa_terminal_rows(),
                             gaffa_hip_ffa_terminal_rows_max(),
                             "FFA terminal rows per item too large");
  check_arg(!plan.tasks.empty(), "FFA executor plan must contain at least one task");
  input_nsamples_ = plan.tasks.front().input_nsamples;
  check_arg(input_nsamples_ > 0, "FFA executor time series must not be empty");
  lds_max_bins_ = plan_max_bins(plan);
  append(plan, 0, 0);
  finish();
}

GpuFfaExecutionPlan::GpuFfaExecutionPlan(std::span<const FfaBatchMember> members,
                                         FfaItemShape shape) {
  level_rows_ = pick_rows(shape.level_rows, gaffa_hip_ffa_level_rows(),
                          gaffa_hip_ffa_level_rows_max(), "FFA level rows per item too large");
  terminal_rows_ = pick_rows(shape.terminal_rows, gaffa_hip_ffa_terminal_rows(),
                             gaffa_hip_ffa_terminal_rows_max(),
                             "FFA terminal rows per item too large");
  check_arg(!members.empty(), "FFA batch must contain at least one series");
  plan_.width_trials = members.front().plan->width_trials;
  for (const FfaBatchMember& m : members)
    if (m.plan != nullptr) lds_max_bins_ = std::max(lds_max_bins_, plan_max_bins(*m.plan));
  for (const FfaBatchMember& m : members) {
    check_arg(m.plan != nullptr && !m.plan->tasks.empty(),
              "FFA executor plan must contain at least one task");
    check_arg(m.plan->width_trials == plan_.width_trials,
              "FFA batch members must share one width-trial list");
    check_arg(m.input_offset % 64 == 0, "FFA batch member offsets must be 64-float aligned");
    const std::size_t n = m.plan->tasks.front().input_nsamples;
    check_arg(n > 0, "FFA executor time series must not be empty");
    input_nsamples_ = std::max<std::size_t>(input_nsamples_, m.input_offset + n);
    append(*m.plan, m.input_offset, m.series);
    plan_.tasks.insert(plan_.tasks.end(), m.plan->tasks.begin(), m.plan->tasks.end());
  }
  finish();
}

void GpuFfaExecutionPlan::append(const FfaSearchPlan& plan, std::uint64_t input_offset,
                                 std::uint32_t series) {
  check_arg(!plan.width_trials.empty(), "FFA search plan width_trials must not be empty");
  check_arg(plan.width_trials.size() <= GAFFA_HIP_FFA_MAX_WIDTHS,
            "GPU FFA supports at most 16 width trials");
  const std::size_t n = plan.tasks.front().input_nsamples;
  const std::uint32_t level_rows = level_rows_;
  const std::uint32_t terminal_rows = terminal_rows_;
  std::map<double, std::uint32_t> group_of;
  for (const FfaSearchTask& t : plan.tasks) {
    validate_task(t, n);
    if (group_of.count(t.downsample_factor) == 0) {
      gaffa_hip_ffa_group g{};
      g.factor = t.downsample_factor;
      g.prepared_nsamples = t.prepared_nsamples;
      g.input_offset = input_offset;
      g.input_nsamples = n;
      if (t.downsample_factor != 1.0) {
        g.prep_offset = prepared_stride_;
        prepared_stride_ += align_up(t.prepared_nsamples);
        max_prepared_ = std::max(max_prepared_, t.prepared_nsamples);
        const std::size_t per_block = gaffa_hip_ffa_prepare_outputs_per_block(t.downsample_factor);
        const std::size_t blocks = (t.prepared_nsamples + per_block - 1) / per_block;
        check_arg(blocks < (1ULL << 31), "FFA prepare grid too large");
        prep_blocks_ = std::max(prep_blocks_, static_cast<std::uint32_t>(blocks));
      }
      group_of[t.downsample_factor] = static_cast<std::uint32_t>(groups_.size());
      groups_.push_back(g);
    }
  }
  std::uint64_t periodogram_row = 0;
  for (const FfaSearchTask& t : plan.tasks) {
    const FfaDetectionPlan det = make_ffa_detection_plan(plan.width_trials, t.bins);
    const float stdnoise = ffa_task_stdnoise(t);
    check_arg(stdnoise > 0.0F && std::isfinite(stdnoise),
              "FFA row detection stdnoise must be finite and > 0");
    gaffa_hip_ffa_task d{};
    d.terminal_levels = 1;
    d.group = group_of[t.downsample_factor];
    d.bins = static_cast<std::uint32_t>(t.bins);
    d.rows = static_cast<std::uint32_t>(t.rows);
    d.rows_eval = static_cast<std::uint32_t>(t.rows_eval);
    d.nwidths = static_cast<std::uint32_t>(det.boxcar_trials.size());
    d.inv_stdnoise = 1.0F / stdnoise;
    d.series = series;
    d.periodogram_row = static_cast<std::uint32_t>(periodogram_row);
    periodogram_row += t.rows_eval;
    check_arg(periodogram_row < (1ULL << 32), "FFA plan has too many evaluated rows");
    for (std::size_t w = 0; w < det.boxcar_trials.size(); ++w) {
      d.width[w] = static_cast<std::uint32_t>(det.boxcar_trials[w].width);
      d.height[w] = det.boxcar_trials[w].height;
      d.baseline[w] = det.boxcar_trials[w].baseline;
    }
    check_arg(det.max_width <= t.bins, "FFA detection width exceeds bins");
    max_bins_ = std::max(max_bins_, d.bins);
    terminal_lds_bytes_ = std::max(
        terminal_lds_bytes_,
        gaffa_hip_ffa_terminal_lds_bytes_rows(d.bins, static_cast<std::uint32_t>(det.max_width),
                                              terminal_rows));
    // rows of a node: two ping-pong buffers in what the fixed LDS leaves
    check_arg(subtree_fixed_lds(lds_max_bins_) < kSubtreeLdsBytes,
              "GPU FFA bins too large for an LDS subtree");
    const std::size_t node_floats = (kSubtreeLdsBytes - subtree_fixed_lds(lds_max_bins_)) / 8;
    const std::size_t lds_rows = std::min(kSubtreeMaxRows, node_floats / t.bins);
    check_arg(lds_rows >= 2, "GPU FFA bins too large for an LDS subtree");
    const std::uint32_t task_index = static_cast<std::uint32_t>(tasks_.size());
    if (t.rows >= 2) {
      std::uint32_t depth = 1;
      while ((t.rows + (std::size_t{1} << depth) - 1) >> depth > lds_rows) ++depth;
      d.depth = depth;
      d.work_offset = work_stride_;
      work_stride_ += align_up(t.rows * t.bins);
      const std::size_t before = subtree_.size();
      collect_nodes(0, d.rows, 0, depth, task_index, subtree_);
      for (std::size_t i = before; i < subtree_.size(); ++i) {
        gaffa_hip_ffa_item& it = subtree_[i];
        check_arg(gaffa_hip_ffa_subtree_lds_bytes(it.b, d.bins) != 0xFFFFFFFFu,
                  "GPU FFA subtree node exceeds the kernel limits");
        subtree_half_ = std::max<std::uint32_t>(subtree_half_, it.b * d.bins);
        const std::uint32_t entries = gaffa_hip_ffa_subtree_merges(it.b, d.bins, nullptr);
        check_arg(entries > 0 || it.b < 2, "GPU FFA subtree merge table unavailable");
        subtree_desc_ = std::max(subtree_desc_, entries);
        // one merge table per (rows, bins), shared by every node of that shape
        auto [pos, fresh] = merge_table_of_.try_emplace(
            (std::uint64_t{it.b} << 32) | d.bins, static_cast<std::uint32_t>(merges_.size()));
        if (fresh) {
          check_arg(merges_.size() + entries < (std::size_t{1} << 32),
                    "GPU FFA subtree merge tables too large");
          merges_.resize(merges_.size() + entries);
          check_arg(gaffa_hip_ffa_subtree_merges(it.b, d.bins, merges_.data() + pos->second) ==
                        entries,
                    "GPU FFA subtree merge table size mismatch");
        }
        it.c = pos->second;
      }
      subtree_lds_bytes_ =
          static_cast<std::uint32_t>(8u * subtree_half_ + subtree_fixed_lds(lds_max_bins_));
      if (by_depth_.size() < depth) {
        by_depth_.resize(depth);
        spans_by_depth_.resize(depth);
      }
      // global passes from depth `depth` up to depth `tl`: three tree levels
      // per pass, two where two or four remain (2 + 2 reads less than
      // 3 + 1); the region bits make every pass read one work region and
      // write the other.  Every node above the LDS depth has two rows or
      // more (the depth is the first whose nodes fit), as a multi-level pass
      // and a two-level detection merge need.  The detection kernel merges
      // the top two levels itself (tl = 2) where that saves a pass: one, four
      // or seven levels above depth 1.
      const auto passes = [](std::uint32_t levels) {
        std::uint32_t n = 0;
        while (levels > 0) {
          levels -= levels == 4 ? 2u : std::min<std::uint32_t>(levels, 3u);
          ++n;
        }
        return n;
      };
      const std::uint32_t tl = depth >= 2 && passes(depth - 2) < passes(depth - 1) ? 2u : 1u;
      d.terminal_levels = tl;
      std::uint32_t regions = 0;  // the subtrees write region 0
      for (std::uint32_t din = depth; din > tl;) {
        const std::uint32_t rem = din - tl;
        const std::uint32_t levels = rem == 4 ? 2u : std::min<std::uint32_t>(rem, 3u);
        const std::uint32_t dout = din - levels;
        regions |= (1u - ((regions >> din) & 1u)) << dout;
        // items of up to level_rows rows inside one depth-dout node, so that
        // each leaf of an item's rows is one short run of rows; the level
        // launch gets the LDS of the largest item within kLevelLdsCap, and
        // items beyond it merge from global memory
        std::vector<gaffa_hip_ffa_item> nodes;
        collect_nodes(0, d.rows, 0, dout, task_index, nodes);
        for (const gaffa_hip_ffa_item& nd : nodes) {
          for (std::uint32_t r = nd.a; r < nd.a + nd.b; r += level_rows) {
            const std::uint32_t count = std::min(level_rows, nd.a + nd.b - r);
            by_depth_[dout].push_back(
                gaffa_hip_ffa_item{task_index, dout | (levels << 16), r, count});
            gaffa_hip_ffa_level_span span{};
            check_arg(level_leaf_rows_.size() + count < (std::size_t{1} << 32),
                      "FFA plan has too many level rows");
            span.row_offset = static_cast<std::uint32_t>(level_leaf_rows_.size());
            level_leaf_rows_.resize(level_leaf_rows_.size() + count);
            const std::uint64_t need = gaffa_hip_ffa_level_item(
                d.rows, d.bins, dout, levels, r, count, &span,
                level_leaf_rows_.data() + span.row_offset);
            check_arg(need != ~std::uint64_t{0}, "FFA level item outside its task");
            spans_by_depth_[dout].push_back(span);
            if (need <= kLevelLdsCap)
              level_lds_bytes_ = std::max(level_lds_bytes_, static_cast<std::uint32_t>(need));
          }
        }
        din = dout;
      }
      d.regions = regions;
      max_depth_ = std::max<std::size_t>(max_depth_, depth);
    }
    for (std::uint32_t r = 0; r < d.rows_eval; r += terminal_rows)
      terminal_.push_back(
          gaffa_hip_ffa_item{task_index, r, std::min(terminal_rows, d.rows_eval - r), 0});
    tasks_.push_back(d);
  }
}

void GpuFfaExecutionPlan::finish() {
  // one prepare pass per series when every downsample group reads the same
  // series and the widest window fits the fused kernel's halo
  double max_factor = 0.0;
  bool one_series = groups_.size() <= 64;
  for (const gaffa_hip_ffa_group& g : groups_) {
    one_series = one_series && g.input_offset == 0 && g.input_nsamples == input_nsamples_;
    if (g.factor != 1.0) max_factor = std::max(max_factor, g.factor);
  }
  const double halo = std::ceil(max_factor) + 2.0;
  prep_fused_halo_ = one_series && max_factor > 0.0 && halo <= 4096.0
                         ? static_cast<std::uint32_t>(halo)
                         : 0u;
  level_offsets_.assign(max_depth_ + 1, 0);
  for (std::size_t l = 0; l <= max_depth_; ++l) {
    level_offsets_[l] = static_cast<std::uint32_t>(level_.size());
    if (l < by_depth_.size()) {
      level_.insert(level_.end(), by_depth_[l].begin(), by_depth_[l].end());
      level_spans_.insert(level_spans_.end(), spans_by_depth_[l].begin(),
                          spans_by_depth_[l].end());
    }
  }
  level_offsets_.push_back(static_cast<std::uint32_t>(level_.size()));
  by_depth_.clear();
  spans_by_depth_.clear();
  check_arg(terminal_.size() < (1ULL << 31) && tasks_.size() < (1ULL << 31),
            "FFA plan has too many work items");
}

std::size_t GpuFfaExecutionPlan::workspace_bytes(std::size_t series) const noexcept {
  return series * (prepared_stride_ + 2 * work_stride_) * sizeof(float);
}

GpuFfaProgram::GpuFfaProgram(const FfaSearchPlan& plan, const GpuFfaProgramOptions& po,
                             const GpuFfaExecutionOptions& eo)
    : exec_plan_(plan), program_(po), exec_(eo) {
  GpuDeviceScope scope(program_.device_id);
  const std::size_t per_series = exec_plan_.workspace_bytes(1);
  if (exec_.series_tile_size != 0) {
    tile_ = exec_.series_tile_size;
  } else {
    tile_ = per_series == 0 ? 1024 : std::max<std::size_t>(1, exec_.workspace_bytes_limit / per_series);
    tile_ = std::min<std::size_t>(tile_, 1024);
  }
  tile_ = std::min<std::size_t>(tile_, 65535);
  if (exec_.workspace_bytes_limit != 0 && exec_plan_.workspace_bytes(tile_) > exec_.workspace_bytes_limit &&
      exec_.series_tile_size != 0) {
    throw std::invalid_argument("FFA workspace exceeds workspace_bytes_limit");
  }
  const auto& g = exec_plan_.groups();
  const auto& t = exec_plan_.tasks();
  d_groups_ = upload(g.data(), g.size());
  d_tasks_ = upload(t.data(), t.size());
  d_subtree_ = upload(exec_plan_.subtree_items().data(), exec_plan_.subtree_items().size());
  d_merges_ = upload(exec_plan_.subtree_merges().data(), exec_plan_.subtree_merges().size());
  d_level_ = upload(exec_plan_.level_items().data(), exec_plan_.level_items().size());
  d_level_spans_ = upload(exec_plan_.level_spans().data(), exec_plan_.level_spans().size());
  d_level_rows_ = upload(exec_plan_.level_leaf_rows().data(), exec_plan_.level_leaf_rows().size());
  d_terminal_ = upload(exec_plan_.terminal_items().data(), exec_plan_.terminal_items().size());
  prepared_ = GpuDeviceBuffer<float>(std::max<std::size_t>(1, tile_ * exec_plan_.prepared_stride()));
  work_ = GpuDeviceBuffer<float>(std::max<std::size_t>(1, 2 * tile_ * exec_plan_.work_stride()));
  const std::size_t cap =
      std::max<std::size_t>(1, exec_.initial_peak_buffer_bytes / sizeof(gaffa_hip_ffa_peak));
  peaks_ = GpuDeviceBuffer<gaffa_hip_ffa_peak>(cap);
  peak_count_ = GpuDeviceBuffer<std::uint32_t>(1);
  desc_.groups = d_groups_.data();
  desc_.ngroups = static_cast<std::uint32_t>(g.size());
  desc_.tasks = d_tasks_.data();
  desc_.ntasks = static_cast<std::uint32_t>(t.size());
  desc_.subtree_items = d_subtree_.data();
  desc_.nsubtree_items = static_cast<std::uint32_t>(exec_plan_.subtree_items().size());
  desc_.subtree_merges = d_merges_.data();
  desc_.level_items = d_level_.data();
  desc_.level_spans = d_level_spans_.data();
  desc_.level_leaf_rows = d_level_rows_.data();
  desc_.level_item_offsets = exec_plan_.level_offsets().data();
  desc_.max_depth = static_cast<std::uint32_t>(exec_plan_.max_depth());
  desc_.terminal_items = d_terminal_.data();
  desc_.nterminal_items = static_cast<std::uint32_t>(exec_plan_.terminal_items().size());
  desc_.max_bins = exec_plan_.max_bins();
  desc_.subtree_lds_bytes = exec_plan_.subtree_lds_bytes();
  desc_.subtree_half_floats = exec_plan_.subtree_half_floats();
  desc_.max_subtree_desc = exec_plan_.max_subtree_desc();
  desc_.terminal_lds_bytes = exec_plan_.terminal_lds_bytes();
  desc_.level_lds_bytes = exec_plan_.level_lds_bytes();
  desc_.prep_blocks = exec_plan_.prep_blocks();
  desc_.prep_fused_halo = exec_plan_.prep_fused_halo();
  desc_.prep_input_nsamples = exec_plan_.input_nsamples();
  desc_.max_prepared_nsamples = exec_plan_.max_prepared();
  desc_.prepared_stride = exec_plan_.prepared_stride();
  desc_.work_stride = exec_plan_.work_stride();
}

std::size_t GpuFfaProgram::workspace_bytes() const noexcept {
  return prepared_.bytes() + work_.bytes() + peaks_.bytes();
}

gaffa_hip_ffa_search_args GpuFfaProgram::args(const float* series, std::size_t nseries,
                                              std::size_t stride, std::uint32_t base, float thr) {
  desc_.level_item_offsets = exec_plan_.level_offsets().data();
  gaffa_hip_ffa_search_args a{};
  a.plan = &desc_;
  a.series = series;
  a.nseries = nseries;
  a.nsamples = exec_plan_.input_nsamples();
  a.series_stride = stride;
  a.series_index_base = base;
  a.snr_threshold = thr;
  a.prepared = prepared_.data();
  a.work = work_.data();
  a.peaks = peaks_.data();
  a.peak_capacity = peaks_.size();
  a.peak_count = peak_count_.data();
  if (periodogram_ != nullptr) {
    a.periodogram = periodogram_ + static_cast<std::size_t>(base) * periodogram_stride_;
    a.periodogram_stride = periodogram_stride_;
    a.periodogram_widths = periodogram_widths_;
  }
  return a;
}

void GpuFfaProgram::set_periodogram(float* base, std::size_t stride, std::uint32_t widths) {
  if (base != nullptr) {
    check_arg(widths >= exec_plan_.plan().width_trials.size() &&
                  widths <= GAFFA_HIP_FFA_MAX_WIDTHS,
              "FFA periodogram needs a slot for every width trial");
    check_arg(stride >= periodogram_rows() * widths, "FFA periodogram stride too small");
  }
  periodogram_ = base;
  periodogram_stride_ = stride;
  periodogram_widths_ = widths;
}

std::size_t GpuFfaProgram::periodogram_rows() const noexcept {
  std::size_t n = 0;
  for (const FfaSearchTask& t : exec_plan_.plan().tasks) n += t.rows_eval;
  return n;
}

void GpuFfaProgram::enqueue_tile(const float* series, std::size_t nseries, std::size_t stride,
                                 std::uint32_t base, float thr, bool reset_count,
                                 const FfaStageMarks* marks) {
  check_arg(nseries > 0 && nseries <= tile_, "FFA tile exceeds the program tile capacity");
  GpuDeviceScope scope(program_.device_id);
  if (reset_count) {
    hip_ok(gaffa_hip_memset(peak_count_.data(), 0, sizeof(std::uint32_t), exec_.stream),
           "hipMemset");
  }
  const gaffa_hip_ffa_search_args a = args(series, nseries, stride, base, thr);
  if (marks == nullptr) {
    hip_ok(gaffa_hip_ffa_search_batch(&a, exec_.stream), "ffa_search_batch");
    return;
  }
  using Stage = int (*)(const gaffa_hip_ffa_search_args*, gaffa_hip_stream);
  static constexpr Stage stages[FfaStageMarks::kStages] = {
      gaffa_hip_ffa_prepare_batch, gaffa_hip_ffa_subtree_batch, gaffa_hip_ffa_levels_batch,
      gaffa_hip_ffa_detect_batch};
  hip_ok(gaffa_hip_event_record(marks->at[0], exec_.stream), "hipEventRecord");
  for (int k = 0; k < FfaStageMarks::kStages; ++k) {
    hip_ok(stages[k](&a, exec_.stream), "ffa_search_batch");
    hip_ok(gaffa_hip_event_record(marks->at[k + 1], exec_.stream), "hipEventRecord");
  }
}

std::uint32_t GpuFfaProgram::read_peak_count() {
  std::uint32_t n = 0;
  copy_to_host(&n, peak_count_.data(), sizeof(n), exec_.stream);
  return n;
}

std::vector<gaffa_hip_ffa_peak> GpuFfaProgram::read_peaks(std::uint32_t count) {
  std::vector<gaffa_hip_ffa_peak> v(std::min<std::size_t>(count, peaks_.size()));
  if (!v.empty()) copy_to_host(v.data(), peaks_.data(), v.size() * sizeof(v[0]), exec_.stream);
  return v;
}

std::size_t GpuFfaProgram::read_peaks_into(std::uint32_t count, gaffa_hip_ffa_peak* dst) {
  const std::size_t n = std::min<std::size_t>(count, peaks_.size());
  if (n != 0) copy_to_host(dst, peaks_.data(), n * sizeof(gaffa_hip_ffa_peak), exec_.stream);
  hip_ok(gaffa_hip_stream_synchronize(exec_.stream), "hipStreamSynchronize");
  return n;
}

void GpuFfaProgram::enqueue_peak_count_copy(std::uint32_t* host_slot) {
  hip_ok(gaffa_hip_memcpy(host_slot, peak_count_.data(), sizeof(std::uint32_t), exec_.stream),
         "hipMemcpyAsync peak count");
}

void GpuFfaProgram::enqueue_peak_copy(std::size_t first, std::size_t count, gaffa_hip_ffa_peak* dst,
                                      gaffa_hip_stream stream) const {
  check_arg(first + count <= peaks_.size(), "peak range exceeds the peak buffer");
  if (count != 0)
    hip_ok(gaffa_hip_memcpy(dst, peaks_.data() + first, count * sizeof(gaffa_hip_ffa_peak), stream),
           "hipMemcpyAsync peaks");
}

void GpuFfaProgram::grow_peaks(std::size_t count) {
  const std::size_t max_count = exec_.max_peak_buffer_bytes / sizeof(gaffa_hip_ffa_peak);
  if (count > max_count) {
    throw std::runtime_error("FFA compact peak buffer exceeded max_peak_buffer_bytes");
  }
  GpuDeviceScope scope(program_.device_id);
  peaks_ = GpuDeviceBuffer<gaffa_hip_ffa_peak>(std::min(max_count, std::max(count, 2 * peaks_.size())));
}

void GpuFfaProgram::transform_rows(const float* series, std::size_t nseries, std::size_t stride,
                                   std::size_t task, float* out) {
  check_arg(nseries > 0 && nseries <= tile_, "FFA tile exceeds the program tile capacity");
  check_arg(task < exec_plan_.tasks().size(), "FFA task index out of range");
  GpuDeviceScope scope(program_.device_id);
  const gaffa_hip_ffa_search_args a = args(series, nseries, stride, 0, 0.0F);
  hip_ok(gaffa_hip_ffa_transform_rows(&a, static_cast<std::uint32_t>(task), out, exec_.stream),
         "ffa_transform_rows");
}

std::vector<FfaBatchPeak> expand_ffa_peaks(const GpuFfaExecutionPlan& plan,
                                           std::span<const gaffa_hip_ffa_peak> peaks) {
  const auto& tasks = plan.plan().tasks;
  // one detection plan per task (tasks index them directly in the loop)
  std::map<std::size_t, FfaDetectionPlan> by_bins;
  std::vector<const FfaDetectionPlan*> det(tasks.size());
  for (std::size_t k = 0; k < tasks.size(); ++k) {
    auto it = by_bins.find(tasks[k].bins);
    if (it == by_bins.end())
      it = by_bins.emplace(tasks[k].bins,
                           make_ffa_detection_plan(plan.plan().width_trials, tasks[k].bins)).first;
    det[k] = &it->second;
  }
  for (const gaffa_hip_ffa_peak& p : peaks) {
    if (p.task >= tasks.size() || (p.phase_width >> 16) >= det[p.task]->boxcar_trials.size()) {
      throw std::logic_error("FFA compact peak refers to an unknown task or width");
    }
  }
  std::vector<FfaBatchPeak> out(peaks.size());
#pragma omp parallel for schedule(static)
  for (std::int64_t i = 0; i < static_cast<std::int64_t>(peaks.size()); ++i) {
    const gaffa_hip_ffa_peak& p = peaks[static_cast<std::size_t>(i)];
    out[static_cast<std::size_t>(i)] = FfaBatchPeak{
        p.series, make_ffa_peak(tasks[p.task], det[p.task]->boxcar_trials[p.phase_width >> 16],
                                p.shift, p.phase_width & 0xFFFFu, p.snr)};
  }
  return out;
}

std::vector<DmPeak> expand_sorted_dm_peaks(const GpuFfaExecutionPlan& plan,
                                           std::span<const gaffa_hip_ffa_peak> peaks,
                                           std::span<const double> dms) {
  std::vector<DmPeak> out;
  expand_sorted_dm_peaks_into(plan, peaks, dms, out);
  return out;
}

std::vector<std::uint32_t> ffa_period_ranks(const FfaSearchPlan& plan) {
  const std::vector<double> periods = ffa_plan_trial_periods(plan);
  if (periods.size() >= (std::size_t{1} << 31)) return {};
  for (double v : periods)
    if (!std::isfinite(v)) return {};
  std::vector<std::uint32_t> order(periods.size());
  for (std::size_t i = 0; i < order.size(); ++i) order[i] = static_cast<std::uint32_t>(i);
  __gnu_parallel::sort(order.begin(), order.end(), [&](std::uint32_t a, std::uint32_t b) {
    return periods[a] < periods[b] || (periods[a] == periods[b] && a < b);
  });
  std::vector<std::uint32_t> rank(periods.size());
  std::uint32_t r = 0;
  for (std::size_t i = 0; i < order.size(); ++i) {
    if (i > 0 && periods[order[i]] != periods[order[i - 1]]) ++r;
    rank[order[i]] = r;
  }
  return rank;
}

namespace {

// Detection plans by bins, one per task (make_ffa_peak's width tables).
std::vector<const FfaDetectionPlan*> detection_plans(const GpuFfaExecutionPlan& plan,
                                                     std::map<std::size_t, FfaDetectionPlan>& by_bins) {
  const auto& tasks = plan.plan().tasks;
  std::vector<const FfaDetectionPlan*> det(tasks.size());
  for (std::size_t k = 0; k < tasks.size(); ++k) {
    auto it = by_bins.find(tasks[k].bins);
    if (it == by_bins.end())
      it = by_bins.emplace(tasks[k].bins,
                           make_ffa_detection_plan(plan.plan().width_trials, tasks[k].bins)).first;
    det[k] = &it->second;
  }
  return det;
}

DmPeak expand_dm_peak(const std::vector<FfaSearchTask>& tasks,
                      const std::vector<const FfaDetectionPlan*>& det, std::span<const double> dms,
                      const gaffa_hip_ffa_peak& p) {
  return DmPeak{dms[p.series], p.series,
                make_ffa_peak(tasks[p.task], det[p.task]->boxcar_trials[p.phase_width >> 16],
                              p.shift, p.phase_width & 0xFFFFu, p.snr)};
}

}  // namespace

void expand_sorted_dm_peaks_into(const GpuFfaExecutionPlan& plan,
                                 std::span<const gaffa_hip_ffa_peak> peaks,
                                 std::span<const double> dms, std::vector<DmPeak>& out,
                                 std::span<const std::uint32_t> period_ranks) {
  std::vector<DmPeakKey> keys(peaks.size());
  dm_peak_keys(plan, peaks, 0, dms, period_ranks, keys.data());
  finish_sorted_dm_peaks(plan, peaks, dms, keys, out);
}

// Sort (key, index) pairs instead of 96-byte records.  The key orders by
// is_better_dm_peak's first criteria: finite peaks first, then S/N
// descending (-0 and +0 share a key because they compare equal), then, for
// finite peaks of a plan with period ranks, period ascending (frequency
// follows the period).  Runs of equal keys are then ordered with the full
// comparator, so the result is exactly std::sort(..., is_better_dm_peak) of
// the expanded peaks.
void dm_peak_keys(const GpuFfaExecutionPlan& plan, std::span<const gaffa_hip_ffa_peak> peaks,
                  std::size_t first, std::span<const double> dms,
                  std::span<const std::uint32_t> period_ranks, DmPeakKey* keys) {
  const auto& tasks = plan.plan().tasks;
  std::map<std::size_t, FfaDetectionPlan> by_bins;
  const std::vector<const FfaDetectionPlan*> det = detection_plans(plan, by_bins);
  // first periodogram row of each task (plan order, rows_eval each)
  std::vector<std::size_t> first_row(tasks.size() + 1, 0);
  for (std::size_t k = 0; k < tasks.size(); ++k) first_row[k + 1] = first_row[k] + tasks[k].rows_eval;
  const bool ranked = !period_ranks.empty() && period_ranks.size() == first_row.back();
  for (const gaffa_hip_ffa_peak& p : peaks) {
    if (p.task >= tasks.size() || (p.phase_width >> 16) >= det[p.task]->boxcar_trials.size() ||
        p.series >= dms.size() || p.shift >= tasks[p.task].rows_eval) {
      throw std::logic_error("FFA compact peak refers to an unknown task, row, width or series");
    }
  }
  const std::int64_t n = static_cast<std::int64_t>(peaks.size());
#pragma omp parallel for schedule(static)
  for (std::int64_t i = 0; i < n; ++i) {
    const gaffa_hip_ffa_peak& c = peaks[static_cast<std::size_t>(i)];
    const FfaPeak e = expand_dm_peak(tasks, det, dms, c).peak;
    const bool finite =
        std::isfinite(e.snr) && std::isfinite(e.period) && std::isfinite(e.frequency);
    std::uint64_t k = finite ? 0 : (1ULL << 63);
    if (std::isnan(c.snr)) {
      k |= 0xFFFFFFFFULL << 31;  // one run, ordered by the comparator
    } else {
      const float snr = c.snr == 0.0F ? 0.0F : c.snr;  // canonical zero
      std::uint32_t bits = 0;
      std::memcpy(&bits, &snr, 4);
      const std::uint32_t ascending = (bits & 0x80000000U) ? ~bits : (bits | 0x80000000U);
      k |= static_cast<std::uint64_t>(~ascending) << 31;  // descending S/N
      if (finite && ranked) k |= period_ranks[first_row[c.task] + c.shift];
    }
    keys[i] = DmPeakKey{k, static_cast<std::uint32_t>(first + static_cast<std::size_t>(i))};
  }
}

void finish_sorted_dm_peaks(const GpuFfaExecutionPlan& plan,
                            std::span<const gaffa_hip_ffa_peak> peaks,
                            std::span<const double> dms, std::vector<DmPeakKey>& keys,
                            std::vector<DmPeak>& out) {
  check_arg(keys.size() == peaks.size(), "one sort key per peak");
  const auto& tasks = plan.plan().tasks;
  std::map<std::size_t, FfaDetectionPlan> by_bins;
  const std::vector<const FfaDetectionPlan*> det = detection_plans(plan, by_bins);
  const std::int64_t n = static_cast<std::int64_t>(peaks.size());
  __gnu_parallel::sort(keys.begin(), keys.end(),
                       [](const DmPeakKey& a, const DmPeakKey& b) { return a.key < b.key; });
  // equal-key runs, found and ordered in parallel (each run is independent)
  std::vector<std::pair<std::size_t, std::size_t>> runs;
#pragma omp parallel
  {
    std::vector<std::pair<std::size_t, std::size_t>> mine;
#pragma omp for schedule(static) nowait
    for (std::int64_t s = 0; s < n; ++s) {
      const std::size_t i = static_cast<std::size_t>(s);
      if (i > 0 && keys[i].key == keys[i - 1].key) continue;  // not a run start
      std::size_t j = i + 1;
      while (j < keys.size() && keys[j].key == keys[i].key) ++j;
      if (j - i > 1) mine.emplace_back(i, j);
    }
#pragma omp critical
    runs.insert(runs.end(), mine.begin(), mine.end());
  }
  const std::int64_t nruns = static_cast<std::int64_t>(runs.size());
#pragma omp parallel for schedule(dynamic, 64)
  for (std::int64_t r = 0; r < nruns; ++r) {
    const auto [i, j] = runs[static_cast<std::size_t>(r)];
    std::vector<std::pair<DmPeak, std::uint32_t>> run;
    run.reserve(j - i);
    for (std::size_t k = i; k < j; ++k)
      run.emplace_back(expand_dm_peak(tasks, det, dms, peaks[keys[k].index]), keys[k].index);
    std::sort(run.begin(), run.end(), [](const auto& a, const auto& b) {
      return is_better_dm_peak(a.first, b.first);
    });
    for (std::size_t k = i; k < j; ++k) keys[k].index = run[k - i].second;
  }
  out.resize(peaks.size());
#pragma omp parallel for schedule(static)
  for (std::int64_t i = 0; i < n; ++i)
    out[static_cast<std::size_t>(i)] =
        expand_dm_peak(tasks, det, dms, peaks[keys[static_cast<std::size_t>(i)].index]);
}

void expand_dm_peaks(const GpuFfaExecutionPlan& plan, std::span<const gaffa_hip_ffa_peak> peaks,
                     std::span<const double> dms, DmPeak* out) {
  const auto& tasks = plan.plan().tasks;
  std::map<std::size_t, FfaDetectionPlan> by_bins;
  const std::vector<const FfaDetectionPlan*> det = detection_plans(plan, by_bins);
  const std::int64_t n = static_cast<std::int64_t>(peaks.size());
#pragma omp parallel for schedule(static)
  for (std::int64_t i = 0; i < n; ++i)
    out[i] = expand_dm_peak(tasks, det, dms, peaks[static_cast<std::size_t>(i)]);
}

void sort_dm_peak_keys(std::span<DmPeakKey> keys) {
  __gnu_parallel::sort(keys.begin(), keys.end(),
                       [](const DmPeakKey& a, const DmPeakKey& b) { return a.key < b.key; });
}

void merge_sorted_dm_peaks(std::span<const DmPeak> expanded, std::span<const std::size_t> run_ends,
                           std::vector<DmPeakKey>& keys, std::vector<DmPeakKey>& scratch,
                           std::vector<DmPeak>& out) {
  check_arg(keys.size() == expanded.size(), "one sort key per peak");
  check_arg(!run_ends.empty() && run_ends.back() == keys.size(), "sorted parts must cover the keys");
  const std::size_t n = keys.size();
  const auto by_key = [](const DmPeakKey& a, const DmPeakKey& b) { return a.key < b.key; };
  if (run_ends.size() > 1) {
    std::vector<std::pair<DmPeakKey*, DmPeakKey*>> parts;
    std::size_t begin = 0;
    for (const std::size_t end : run_ends) {
      check_arg(end >= begin, "sorted parts must be in order");
      if (end > begin) parts.emplace_back(keys.data() + begin, keys.data() + end);
      begin = end;
    }
    scratch.resize(n);
    __gnu_parallel::multiway_merge(parts.begin(), parts.end(), scratch.begin(), n, by_key);
    keys.swap(scratch);
  }
  // equal-key runs, found and ordered in parallel (each run is independent)
  const std::int64_t sn = static_cast<std::int64_t>(n);
  std::vector<std::pair<std::size_t, std::size_t>> runs;
#pragma omp parallel
  {
    std::vector<std::pair<std::size_t, std::size_t>> mine;
#pragma omp for schedule(static) nowait
    for (std::int64_t s = 0; s < sn; ++s) {
      const std::size_t i = static_cast<std::size_t>(s);
      if (i > 0 && keys[i].key == keys[i - 1].key) continue;  // not a run start
      std::size_t j = i + 1;
      while (j < n && keys[j].key == keys[i].key) ++j;
      if (j - i > 1) mine.emplace_back(i, j);
    }
#pragma omp critical
    runs.insert(runs.end(), mine.begin(), mine.end());
  }
  const std::int64_t nruns = static_cast<std::int64_t>(runs.size());
#pragma omp parallel for schedule(dynamic, 64)
  for (std::int64_t r = 0; r < nruns; ++r) {
    const auto [i, j] = runs[static_cast<std::size_t>(r)];
    std::sort(keys.begin() + static_cast<std::ptrdiff_t>(i),
              keys.begin() + static_cast<std::ptrdiff_t>(j),
              [&](const DmPeakKey& a, const DmPeakKey& b) {
                return is_better_dm_peak(expanded[a.index], expanded[b.index]);
              });
  }
  out.resize(n);
#pragma omp parallel for schedule(static)
  for (std::int64_t i = 0; i < sn; ++i)
    out[static_cast<std::size_t>(i)] = expanded[keys[static_cast<std::size_t>(i)].index];
}

FfaBatchSearchResult run_ffa_batch_gpu(GpuFfaProgram& program, GpuTimeSeriesBatchView batch,
                                       const FfaSearchOptions& options) {
  check_arg(std::isfinite(options.snr_threshold), "FFA search S/N threshold must be finite");
  check_arg(batch.data != nullptr, "FFA batch data must not be null");
  check_arg(batch.nseries > 0, "FFA batch must contain at least one series");
  check_arg(batch.nsamples == program.execution_plan().input_nsamples(),
            "FFA executor task input_nsamples must match time series size");
  const std::size_t stride = batch.stride == 0 ? batch.nsamples : batch.stride;
  check_arg(stride >= batch.nsamples, "FFA batch stride must be >= nsamples");
  FfaBatchSearchResult result;
  const std::size_t tile = program.tile_capacity();
  for (std::size_t s0 = 0; s0 < batch.nseries; s0 += tile) {
    const std::size_t ns = std::min(tile, batch.nseries - s0);
    const float* base = batch.data + s0 * stride;
    std::uint32_t count = 0;
    for (;;) {
      program.enqueue_tile(base, ns, stride, static_cast<std::uint32_t>(s0), options.snr_threshold,
                           true);
      count = program.read_peak_count();
      if (count <= program.peak_capacity()) break;
      program.grow_peaks(count);  // repeat the tile with room for every peak
    }
    const std::vector<gaffa_hip_ffa_peak> compact = program.read_peaks(count);
    std::vector<FfaBatchPeak> expanded = expand_ffa_peaks(program.execution_plan(), compact);
    result.peaks.insert(result.peaks.end(), expanded.begin(), expanded.end());
  }
  if (options.max_peaks != 0) {
    std::map<std::size_t, std::size_t> per_series;
    for (const auto& p : result.peaks) {
      if (++per_series[p.series_index] > options.max_peaks) {
        throw std::runtime_error("FFA detection peak count exceeded max_peaks safety guard");
      }
    }
  }
  __gnu_parallel::sort(result.peaks.begin(), result.peaks.end(),
            [](const FfaBatchPeak& a, const FfaBatchPeak& b) {
              if (a.series_index != b.series_index) return a.series_index < b.series_index;
              return is_better_ffa_peak(a.peak, b.peak);
            });
  return result;
}

FfaSearchResult search_ffa_gpu(GpuFfaProgram& program, GpuSpan<const float> ts,
                               const FfaSearchOptions& options) {
  GpuTimeSeriesBatchView batch{ts.data, 1, ts.count, ts.count, ts.device_id};
  FfaBatchSearchResult r = run_ffa_batch_gpu(program, batch, options);
  FfaSearchResult out;
  out.peaks.reserve(r.peaks.size());
  for (const auto& p : r.peaks) out.peaks.push_back(p.peak);
  return out;
}

FfaSearchResult search_ffa_gpu_host(std::span<const float> ts, const FfaSearchPlan& plan,
                                    const FfaSearchOptions& options, int device_id) {
  check_arg(std::isfinite(options.snr_threshold), "FFA search S/N threshold must be finite");
  check_arg(!plan.width_trials.empty(), "FFA search plan width_trials must not be empty");
  check_arg(!ts.empty(), "FFA executor time series must not be empty");
  check_arg(!plan.tasks.empty(), "FFA executor plan must contain at least one task");
  for (const FfaSearchTask& t : plan.tasks) validate_task(t, ts.size());
  GpuDeviceScope scope(device_id);
  GpuDeviceBuffer<float> d = upload(ts.data(), ts.size());
  GpuFfaExecutionOptions eo;
  eo.series_tile_size = 1;
  GpuFfaProgram program(plan, GpuFfaProgramOptions{device_id}, eo);
  return search_ffa_gpu(program, GpuSpan<const float>{d.data(), d.size(), device_id}, options);
}

bool is_better_dm_peak(const DmPeak& a, const DmPeak& b) {
  if (is_better_ffa_peak(a.peak, b.peak)) return true;
  if (is_better_ffa_peak(b.peak, a.peak)) return false;
  if (a.dm_index != b.dm_index) return a.dm_index < b.dm_index;
  return a.dm < b.dm;
}

DmSearchResult find_dm_peaks_gpu(GpuDedispersedView<std::uint32_t> plane,
                                 std::span<const double> dms, double tsamp,
                                 const DmSearchOptions& options, int device_id) {
  check_arg(plane.shape.ndm != 0, "DM search dedispersed ndm must be > 0");
  check_arg(plane.shape.nsamples != 0, "DM search dedispersed nsamples must be > 0");
  check_arg(tsamp > 0.0 && std::isfinite(tsamp), "DM search tsamp must be finite and > 0");
  check_arg(dms.size() == plane.shape.ndm, "DM search dms size must match ndm");
  for (double dm : dms) check_arg(std::isfinite(dm), "DM search dms must be finite");
  check_arg(std::isfinite(options.snr_threshold), "DM search S/N threshold must be finite");
  const std::size_t stride = plane.stride == 0 ? plane.shape.nsamples : plane.stride;
  const FfaSearchPlan plan = make_riptide_ffa_plan(plane.shape.nsamples, tsamp, options.plan);
  GpuDeviceScope scope(device_id);
  GpuPreprocessExecutionOptions pe;
  pe.series_tile_size = std::min<std::size_t>(plane.shape.ndm, 1024);
  PreprocessPlan pp = options.preprocess;
  const bool convert_only = pp.steps.empty();
  if (!convert_only) {
    GpuPreprocessProgram prep(pp, GpuPreprocessProgramOptions{device_id}, pe);
    prep.run(plane.data, plane.shape.ndm, plane.shape.nsamples, stride, true);
    prep.synchronize();
  } else {
    gaffa_hip_preprocess_args a{};
    a.data = plane.data;
    a.nseries = plane.shape.ndm;
    a.nsamples = plane.shape.nsamples;
    a.stride = stride;
    a.input_is_u32 = 1;
    GpuDeviceBuffer<std::uint8_t> ws(
        gaffa_hip_preprocess_workspace_bytes(plane.shape.ndm, plane.shape.nsamples, 1));
    GpuDeviceBuffer<std::int32_t> st(plane.shape.ndm);
    hip_ok(gaffa_hip_memset(st.data(), 0, st.bytes(), nullptr), "hipMemset");
    a.workspace = ws.data();
    a.workspace_bytes = ws.bytes();
    a.status = st.data();
    a.scrunch_factor = 1;
    a.median_window = 1;
    hip_ok(gaffa_hip_detrend_normalise_batch(&a, nullptr), "convert_uint32_to_float");
    hip_ok(gaffa_hip_stream_synchronize(nullptr), "hipStreamSynchronize");
  }
  GpuFfaProgram program(plan, GpuFfaProgramOptions{device_id});
  const FfaSearchOptions fo{options.snr_threshold, options.max_peaks};
  GpuTimeSeriesBatchView batch{reinterpret_cast<const float*>(plane.data), plane.shape.ndm,
                               plane.shape.nsamples, stride, device_id};
  const FfaBatchSearchResult r = run_ffa_batch_gpu(program, batch, fo);
  DmSearchResult out;
  out.peaks.reserve(r.peaks.size());
  for (const auto& p : r.peaks) out.peaks.push_back(DmPeak{dms[p.series_index], p.series_index, p.peak});
  __gnu_parallel::sort(out.peaks.begin(), out.peaks.end(), is_better_dm_peak);
  return out;
}

}  // namespace gaffa
