// Batched FFA search over series of different lengths, each with its own
// plan (no gaffa counterpart: gaffa's executor runs one plan over equal-length
// series, ffa_cuda_executor.cuh:385-531, and a caller loops over plans).
//
// The series are packed end to end in one device row (64-float aligned
// offsets).  Consecutive series are grouped into tiles; a tile is one
// GpuFfaExecutionPlan built from its members' plans, so one prepare,
// subtree, level and detect launch sequence covers the whole tile whatever
// its members' lengths and row counts.  Tiles share one workspace and one
// compact peak buffer; the launch sequence of all tiles can be captured once
// in a HIP graph and replayed.  The tile size and the rows per level and
// terminal work item are chosen by timing candidates on the actual batch
// (autotune()); the choice is cached per batch shape for the process.
#pragma once

#include "gaffa/ffa_gpu.h"

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

namespace gaffa {

struct GpuFfaBatchedOptions {
  int device_id = 0;
  std::size_t series_per_tile = 16;
  std::size_t workspace_bytes_limit = std::size_t{32} << 30;
  std::size_t initial_peak_buffer_bytes = std::size_t{64} << 20;
  std::size_t max_peak_buffer_bytes = std::size_t{1} << 30;
  bool use_graph = true;
  FfaItemShape item_shape{};  // rows per level / terminal item (0 = defaults)
};

// What autotune() chose, the device time of that configuration and whether
// it came from the cache of an earlier batch of the same shape.
struct FfaBatchTuning {
  std::size_t series_per_tile = 0;
  FfaItemShape item_shape{};
  float gpu_ms = 0.0F;
  bool cached = false;
};

struct FfaBatchTiming {
  float gpu_ms = 0.0F;        // events around the launch sequence (or graph)
  std::size_t tiles = 0;
  std::size_t launches = 0;   // kernel launches per run
};

class GpuFfaBatchedSearch {
 public:
  // plans[i] searches series i (its length is plans[i].tasks[0].input_nsamples).
  GpuFfaBatchedSearch(std::vector<FfaSearchPlan> plans, const GpuFfaBatchedOptions& options = {});
  ~GpuFfaBatchedSearch();
  GpuFfaBatchedSearch(const GpuFfaBatchedSearch&) = delete;
  GpuFfaBatchedSearch& operator=(const GpuFfaBatchedSearch&) = delete;

  std::size_t nseries() const noexcept { return plans_.size(); }
  const std::vector<std::uint64_t>& offsets() const noexcept { return offsets_; }
  const std::vector<std::uint64_t>& lengths() const noexcept { return lengths_; }
  std::size_t row_length() const noexcept { return row_length_; }
  std::size_t series_per_tile() const noexcept { return series_per_tile_; }
  const FfaSearchPlan& plan(std::size_t i) const { return plans_.at(i); }
  // Device row holding the packed series (owned; fill with upload() or a
  // generator writing to row()).
  float* row() noexcept { return row_.data(); }
  void upload(const float* host_row, std::size_t count);
  gaffa_hip_stream stream() const noexcept { return stream_.get(); }

  // Regroup into tiles of at most n series (drops a captured graph).
  void set_series_per_tile(std::size_t n);
  // Rows per level / terminal work item of the following runs (drops a
  // captured graph).
  void set_item_shape(FfaItemShape shape);
  const FfaItemShape& item_shape() const noexcept { return opt_.item_shape; }
  // Tune the tile size over `tile_candidates`, then the rows per level item
  // over `level_rows`, then the rows per terminal item over `terminal_rows`
  // (one knob at a time, each timed on the actual batch, fastest kept), and
  // cache the result under the batch's shape: a later batch of the same
  // shape takes it without timing.  Returns the tile size.
  std::size_t autotune(const std::vector<std::size_t>& tile_candidates,
                       const std::vector<std::uint32_t>& level_rows = {8, 16, 32},
                       const std::vector<std::uint32_t>& terminal_rows = {16, 32});
  const FfaBatchTuning& tuning() const noexcept { return tuning_; }

  // Search every series: peaks sorted by (series, is_better_ffa_peak).
  FfaBatchSearchResult run(float snr_threshold, FfaBatchTiming* timing = nullptr);

  // Baseline: the same searches as one launch sequence per series (the
  // per-plan loop a gaffa caller writes), GPU time in timing->gpu_ms.
  FfaBatchSearchResult run_per_series(float snr_threshold, FfaBatchTiming* timing = nullptr);

 private:
  struct Tile;
  void build_tiles();
  void enqueue(float snr_threshold);
  std::uint32_t read_count();
  FfaBatchSearchResult collect(std::uint32_t count) const;

  GpuFfaBatchedOptions opt_;
  std::vector<FfaSearchPlan> plans_;
  std::vector<std::uint64_t> offsets_, lengths_;
  std::size_t row_length_ = 0;
  std::size_t series_per_tile_ = 0;
  GpuStream stream_;
  GpuDeviceBuffer<float> row_;
  std::vector<std::unique_ptr<Tile>> tiles_;
  std::vector<int> tile_of_;  // series -> tile
  GpuDeviceBuffer<float> prepared_, work_;
  GpuDeviceBuffer<gaffa_hip_ffa_peak> peaks_;
  GpuDeviceBuffer<std::uint32_t> count_;
  gaffa_hip_graph graph_ = nullptr;
  float graph_threshold_ = 0.0F;
  FfaBatchTuning tuning_{};
  gaffa_hip_event ev_[2] = {nullptr, nullptr};
};

}  // namespace gaffa
