// GPU FFA search program (the HIP counterpart of include/gaffa/ffa_cuda.h:
// CudaFfaExecutionPlan / CudaFfaProgram / run_ffa_batch_cuda /
// search_ffa_cuda) and the batched DM search built on it
// (include/gaffa/dm_search.h's DmPeak and find_dm_peaks_cpu semantics).
#pragma once

#include "gaffa/dedispersion.h"
#include "gaffa/ffa_plan.h"
#include "gaffa/gpu_memory.h"
#include "gaffa/preprocessing.h"

#include <cstddef>
#include <cstdint>
#include <map>
#include <span>
#include <vector>

namespace gaffa {

struct GpuFfaProgramOptions {
  int device_id = 0;
};

struct GpuFfaExecutionOptions {
  // Series processed per launch sequence; 0 picks the largest tile whose
  // workspace fits workspace_bytes_limit.
  std::size_t series_tile_size = 0;
  std::size_t workspace_bytes_limit = 16ULL << 30;
  std::size_t initial_peak_buffer_bytes = 64ULL << 20;
  std::size_t max_peak_buffer_bytes = 1ULL << 30;
  gaffa_hip_stream stream = nullptr;
};

struct FfaBatchPeak {
  std::size_t series_index = 0;
  FfaPeak peak{};
};

struct FfaBatchSearchResult {
  std::vector<FfaBatchPeak> peaks;
};

// Events bracketing the four launch stages of one FFA tile.
struct FfaStageMarks {
  static constexpr int kStages = 4;  // prepare, subtree, levels, detect
  gaffa_hip_event at[kStages + 1] = {};
};

// Host-built device layout of one FfaSearchPlan: prepare groups keyed by
// downsample factor, per-task LDS depth, and the work-item lists of the
// subtree / level / terminal launches (see gaffa_hip.h).
// One series of a heterogeneous batch: its plan, the float offset of its
// samples in the batch row (64-float aligned) and its series index.
struct FfaBatchMember {
  const FfaSearchPlan* plan = nullptr;
  std::uint64_t input_offset = 0;
  std::uint32_t series = 0;
};

// Rows per level and per terminal work item (0 = the kernels' defaults,
// gaffa_hip_ffa_level_rows() / gaffa_hip_ffa_terminal_rows()); the batched
// search's autotuner picks them per input shape.  Results do not depend on
// them.
struct FfaItemShape {
  std::uint32_t level_rows = 0;
  std::uint32_t terminal_rows = 0;
};

class GpuFfaExecutionPlan {
 public:
  explicit GpuFfaExecutionPlan(const FfaSearchPlan& plan, FfaItemShape shape = {});
  // Several series with their own lengths and plans laid end to end in one
  // row: every group reads its member's offset, every task reports its
  // member's series index, and plan().tasks is the concatenation of the
  // members' tasks (indexed by the compact peaks' task field).
  explicit GpuFfaExecutionPlan(std::span<const FfaBatchMember> members, FfaItemShape shape = {});

  const FfaSearchPlan& plan() const noexcept { return plan_; }
  std::size_t input_nsamples() const noexcept { return input_nsamples_; }
  std::size_t prepared_stride() const noexcept { return prepared_stride_; }
  std::size_t work_stride() const noexcept { return work_stride_; }
  std::size_t max_depth() const noexcept { return max_depth_; }
  const std::vector<gaffa_hip_ffa_group>& groups() const noexcept { return groups_; }
  const std::vector<gaffa_hip_ffa_task>& tasks() const noexcept { return tasks_; }
  const std::vector<gaffa_hip_ffa_item>& subtree_items() const noexcept { return subtree_; }
  // the subtree items' merge tables (item.c = first entry)
  const std::vector<gaffa_hip_ffa_merge>& subtree_merges() const noexcept { return merges_; }
  const std::vector<gaffa_hip_ffa_item>& level_items() const noexcept { return level_; }
  const std::vector<gaffa_hip_ffa_level_span>& level_spans() const noexcept { return level_spans_; }
  const std::vector<gaffa_hip_ffa_level_row>& level_leaf_rows() const noexcept {
    return level_leaf_rows_;
  }
  const std::vector<std::uint32_t>& level_offsets() const noexcept { return level_offsets_; }
  const std::vector<gaffa_hip_ffa_item>& terminal_items() const noexcept { return terminal_; }
  std::uint32_t subtree_lds_bytes() const noexcept { return subtree_lds_bytes_; }
  std::uint32_t subtree_half_floats() const noexcept { return subtree_half_; }
  std::uint32_t max_subtree_desc() const noexcept { return subtree_desc_; }
  std::uint32_t terminal_lds_bytes() const noexcept { return terminal_lds_bytes_; }
  std::uint32_t level_lds_bytes() const noexcept { return level_lds_bytes_; }
  std::uint32_t prep_blocks() const noexcept { return prep_blocks_; }
  // > 0 when one pass per series serves every downsample group (all read
  // the one series): samples a window reaches past its start.
  std::uint32_t prep_fused_halo() const noexcept { return prep_fused_halo_; }
  std::uint32_t max_bins() const noexcept { return max_bins_; }
  std::size_t max_prepared() const noexcept { return max_prepared_; }
  // Bytes of device workspace for a tile of `series` series.
  std::size_t workspace_bytes(std::size_t series) const noexcept;

 private:
  void append(const FfaSearchPlan& plan, std::uint64_t input_offset, std::uint32_t series);
  void finish();
  std::vector<std::vector<gaffa_hip_ffa_item>> by_depth_;
  std::vector<std::vector<gaffa_hip_ffa_level_span>> spans_by_depth_;
  std::vector<gaffa_hip_ffa_level_span> level_spans_;   // parallel to level_
  std::vector<gaffa_hip_ffa_level_row> level_leaf_rows_;
  FfaSearchPlan plan_;
  std::size_t input_nsamples_ = 0;
  std::size_t prepared_stride_ = 0;
  std::size_t work_stride_ = 0;
  std::size_t max_depth_ = 0;
  std::size_t max_prepared_ = 0;
  std::size_t lds_max_bins_ = 0;  // bins of the widest task of every plan appended
  std::uint32_t level_rows_ = 0, terminal_rows_ = 0;  // rows per level / terminal item
  std::uint32_t subtree_lds_bytes_ = 0;
  std::uint32_t subtree_half_ = 0;
  std::uint32_t subtree_desc_ = 0;
  std::uint32_t terminal_lds_bytes_ = 0;
  std::uint32_t level_lds_bytes_ = 0;
  std::uint32_t prep_blocks_ = 0;
  std::uint32_t prep_fused_halo_ = 0;
  std::uint32_t max_bins_ = 0;
  std::vector<gaffa_hip_ffa_group> groups_;
  std::vector<gaffa_hip_ffa_task> tasks_;
  std::vector<gaffa_hip_ffa_item> subtree_, level_, terminal_;
  std::vector<gaffa_hip_ffa_merge> merges_;
  std::map<std::uint64_t, std::uint32_t> merge_table_of_;  // rows << 32 | bins -> first entry
  std::vector<std::uint32_t> level_offsets_;
};

// Owns device metadata, workspace and the compact peak buffer for one plan on
// one device.  Move-only, not thread-safe, one active run at a time
// (ffa_cuda.h:141-182).
class GpuFfaProgram {
 public:
  explicit GpuFfaProgram(const FfaSearchPlan& plan, const GpuFfaProgramOptions& program_options = {},
                         const GpuFfaExecutionOptions& execution_options = {});
  GpuFfaProgram(GpuFfaProgram&&) noexcept = default;
  GpuFfaProgram& operator=(GpuFfaProgram&&) noexcept = default;

  const GpuFfaExecutionPlan& execution_plan() const noexcept { return exec_plan_; }
  int device_id() const noexcept { return program_.device_id; }
  std::size_t tile_capacity() const noexcept { return tile_; }
  std::size_t workspace_bytes() const noexcept;

  // Enqueue the whole search of one tile (<= tile_capacity series) without
  // waiting; peaks land in the device peak buffer.  Used by the benchmark
  // and the graph path; run_ffa_batch_gpu() wraps it with the host side.
  // With `marks`, its five events are recorded on the program stream before
  // the prepare launch and after the prepare, subtree, level and detection
  // stages.
  void enqueue_tile(const float* series, std::size_t nseries, std::size_t stride,
                    std::uint32_t series_index_base, float snr_threshold, bool reset_count,
                    const FfaStageMarks* marks = nullptr);
  std::uint32_t read_peak_count();
  std::vector<gaffa_hip_ffa_peak> read_peaks(std::uint32_t count);
  // Copies the first min(count, capacity) compact peaks to `dst` (host,
  // ideally pinned) and waits; returns how many.
  std::size_t read_peaks_into(std::uint32_t count, gaffa_hip_ffa_peak* dst);
  std::size_t peak_capacity() const noexcept { return peaks_.size(); }
  // The device compact peaks of the last search, in arrival order.
  const gaffa_hip_ffa_peak* peak_data() const noexcept { return peaks_.data(); }
  // Enqueues on the program stream a copy of the running peak count into
  // `host_slot` (pinned host memory); valid once the stream reaches it.
  void enqueue_peak_count_copy(std::uint32_t* host_slot);
  // Enqueues on `stream` the copy of compact peaks [first, first + count)
  // (within the capacity) into `dst` (pinned host memory).
  void enqueue_peak_copy(std::size_t first, std::size_t count, gaffa_hip_ffa_peak* dst,
                         gaffa_hip_stream stream) const;
  void grow_peaks(std::size_t count);
  gaffa_hip_stream stream() const noexcept { return exec_.stream; }
  // Dense periodogram output for the following tiles: series i of the
  // whole batch (series_index_base + tile index) writes the S/N of every
  // evaluated (row, width) at base + i * stride + row * widths + width, rows
  // in plan order (periodogram_rows() of them).  nullptr switches it off.
  void set_periodogram(float* base, std::size_t stride, std::uint32_t widths);
  std::size_t periodogram_rows() const noexcept;
  // Transform-only access for testing: rows_eval final rows of one task.
  void transform_rows(const float* series, std::size_t nseries, std::size_t stride,
                      std::size_t task, float* out);

 private:
  gaffa_hip_ffa_search_args args(const float* series, std::size_t nseries, std::size_t stride,
                                 std::uint32_t base, float thr);
  GpuFfaExecutionPlan exec_plan_;
  GpuFfaProgramOptions program_;
  GpuFfaExecutionOptions exec_;
  std::size_t tile_ = 0;
  GpuDeviceBuffer<gaffa_hip_ffa_group> d_groups_;
  GpuDeviceBuffer<gaffa_hip_ffa_task> d_tasks_;
  GpuDeviceBuffer<gaffa_hip_ffa_item> d_subtree_, d_level_, d_terminal_;
  GpuDeviceBuffer<gaffa_hip_ffa_level_span> d_level_spans_;
  GpuDeviceBuffer<gaffa_hip_ffa_level_row> d_level_rows_;
  GpuDeviceBuffer<gaffa_hip_ffa_merge> d_merges_;
  GpuDeviceBuffer<float> prepared_, work_;
  GpuDeviceBuffer<gaffa_hip_ffa_peak> peaks_;
  GpuDeviceBuffer<std::uint32_t> peak_count_;
  gaffa_hip_ffa_plan_desc desc_{};
  float* periodogram_ = nullptr;
  std::size_t periodogram_stride_ = 0;
  std::uint32_t periodogram_widths_ = 0;
};

// Expand compact device peaks into FfaBatchPeak records (make_ffa_peak).
std::vector<FfaBatchPeak> expand_ffa_peaks(const GpuFfaExecutionPlan& plan,
                                           std::span<const gaffa_hip_ffa_peak> peaks);

// Searches a dense device batch [nseries][stride] of preprocessed series
// with the program's plan; returns peaks sorted by (series, peak order).
FfaBatchSearchResult run_ffa_batch_gpu(GpuFfaProgram& program, GpuTimeSeriesBatchView batch,
                                       const FfaSearchOptions& options = {});

// Single-series counterpart of search_ffa_cpu (device input).
FfaSearchResult search_ffa_gpu(GpuFfaProgram& program, GpuSpan<const float> time_series,
                               const FfaSearchOptions& options = {});

// One-shot host convenience: upload, search, return (ffa_bindings.cpp:124-163).
FfaSearchResult search_ffa_gpu_host(std::span<const float> time_series, const FfaSearchPlan& plan,
                                    const FfaSearchOptions& options = {}, int device_id = 0);

// ------------------------------------------------------------ DM search --
struct DmPeak {
  double dm = 0.0;
  std::size_t dm_index = 0;
  FfaPeak peak{};
};

struct DmSearchOptions {
  RiptideFfaPlanOptions plan{};
  PreprocessPlan preprocess{};
  float snr_threshold = 6.0F;
  std::size_t max_peaks = 0;
};

struct DmSearchResult {
  std::vector<DmPeak> peaks;
};

// Expands compact peaks (series = DM index into `dms`) and returns them in
// is_better_dm_peak order (dm_search.cpp:142-199), sorting keys rather than
// records.
std::vector<DmPeak> expand_sorted_dm_peaks(const GpuFfaExecutionPlan& plan,
                                           std::span<const gaffa_hip_ffa_peak> peaks,
                                           std::span<const double> dms);
// The same into `out`, reusing its storage (no allocation or first-touch
// cost when it already holds as many peaks, as in repeated searches).
// `period_ranks` (ffa_period_ranks of the plan, optional) lets the sort key
// carry the period order too, so that far fewer equal-key runs need the full
// comparator.
void expand_sorted_dm_peaks_into(const GpuFfaExecutionPlan& plan,
                                 std::span<const gaffa_hip_ffa_peak> peaks,
                                 std::span<const double> dms, std::vector<DmPeak>& out,
                                 std::span<const std::uint32_t> period_ranks = {});
// The two halves of expand_sorted_dm_peaks_into, so that the keys of one
// part of the peaks can be computed while the device still searches the
// rest: dm_peak_keys fills keys[i] for peaks[i] (index = first + i) and
// validates them; finish_sorted_dm_peaks sorts every key and expands.
struct DmPeakKey {
  std::uint64_t key;
  std::uint32_t index;
};
void dm_peak_keys(const GpuFfaExecutionPlan& plan, std::span<const gaffa_hip_ffa_peak> peaks,
                  std::size_t first, std::span<const double> dms,
                  std::span<const std::uint32_t> period_ranks, DmPeakKey* keys);
void finish_sorted_dm_peaks(const GpuFfaExecutionPlan& plan,
                            std::span<const gaffa_hip_ffa_peak> peaks,
                            std::span<const double> dms, std::vector<DmPeakKey>& keys,
                            std::vector<DmPeak>& out);
// The same order built part by part: expand_dm_peaks fills out[i] for
// peaks[i]; merge_sorted_dm_peaks takes keys already sorted within each part
// (part k ends at run_ends[k]) and `expanded` indexed like the keys, merges
// the parts, orders equal-key runs with is_better_dm_peak and gathers the
// records into `out`.  `scratch` is working memory.
void expand_dm_peaks(const GpuFfaExecutionPlan& plan, std::span<const gaffa_hip_ffa_peak> peaks,
                     std::span<const double> dms, DmPeak* out);
void sort_dm_peak_keys(std::span<DmPeakKey> keys);
void merge_sorted_dm_peaks(std::span<const DmPeak> expanded, std::span<const std::size_t> run_ends,
                           std::vector<DmPeakKey>& keys, std::vector<DmPeakKey>& scratch,
                           std::vector<DmPeak>& out);
// Dense rank of every evaluated (task, shift) by trial period (equal periods
// share a rank), indexed by the task's first periodogram row + shift; empty
// when a period is not finite or there are 2^31 rows or more.
std::vector<std::uint32_t> ffa_period_ranks(const FfaSearchPlan& plan);

bool is_better_dm_peak(const DmPeak& lhs, const DmPeak& rhs);

// GPU counterpart of find_dm_peaks_cpu (dm_search.cpp:142-199): the device
// plane (uint32 sums) is converted, preprocessed in place and searched.
// The plane's contents are consumed (it holds normalised floats afterwards).
DmSearchResult find_dm_peaks_gpu(GpuDedispersedView<std::uint32_t> plane,
                                 std::span<const double> dms, double tsamp,
                                 const DmSearchOptions& options, int device_id = 0);

}  // namespace gaffa
