// Fast folding algorithm (FFA) transform and boxcar S/N detection on gfx950.
//
// The transform follows gaffa's recursion exactly (ffa.cpp:110-251): a node of
// R rows splits into head = floor(R/2) rows and tail = R - head rows, and
// output row s of the node is
//     out[s][i] = head[h][i] + tail[t][(i + s - t) mod bins],
//     h = (size_t)(fl32((H-1)/(R-1)) * s + 0.5f),  t likewise with T,
// so every FP32 add matches the CPU's and the transform is bit-exact.
//
// Execution (all series of a batch share one host-built plan):
//   prepare   one thread per prepared sample: riptide weighted-sum downsample
//             with the CPU's serial FP32 order (time_series.cpp:140-174)
//   subtree   one workgroup per (node at the task's LDS depth D, series): the
//             node's rows are staged in LDS and every level below it runs
//             there, ping-ponging between two LDS buffers
//   level     one launch per depth D-1 .. 1: global merges through L2
//   terminal  one workgroup per (64 final rows, series): the root merge into
//             LDS (rows padded with a circular extension of max_width bins),
//             then one lane per (row, width) runs the serial FP32 prefix twice
//             (P[phi] and P[phi+w]) so every boxcar sum equals the CPU's
//             circular-prefix difference bit for bit (ffa_detection.cpp:
//             126-239); significant peaks are appended with one atomic each.
#include "common.hpp"

#include <cmath>

namespace gaffa_hip {
namespace {

constexpr uint32_t kLevelThreads = 256;
constexpr uint32_t kSubtreeThreads = 512;
constexpr uint32_t kTerminalThreads = 256;
constexpr uint32_t kTerminalRows = 64;
constexpr uint32_t kMaxLdsBytes = 160 * 1024;

// ffa.cpp:133-143 in FP32, no contraction (the library is built with
// -ffp-contract=off).
__device__ __forceinline__ uint32_t pick_row(uint32_t child_rows, uint32_t rows,
                                             uint32_t s) {
  const float scale = static_cast<float>(child_rows - 1) / static_cast<float>(rows - 1);
  return static_cast<uint32_t>(scale * static_cast<float>(s) + 0.5f);
}

// Kernel-side view of gaffa_hip_ffa_search_args: the plan descriptor is
// copied by value (it lives in host memory on the C-ABI side).
struct KArgs {
  gaffa_hip_ffa_plan_desc plan;
  const float* series;
  uint64_t nseries;
  uint64_t nsamples;
  uint64_t series_stride;
  uint32_t series_index_base;
  float snr_threshold;
  float* prepared;
  float* work;
  gaffa_hip_ffa_peak* peaks;
  uint64_t peak_capacity;
  uint32_t* peak_count;
};

KArgs kargs(const gaffa_hip_ffa_search_args* a) {
  KArgs k;
  k.plan = *a->plan;
  k.plan.level_item_offsets = nullptr;  // host-only
  k.series = a->series;
  k.nseries = a->nseries;
  k.nsamples = a->nsamples;
  k.series_stride = a->series_stride;
  k.series_index_base = a->series_index_base;
  k.snr_threshold = a->snr_threshold;
  k.prepared = a->prepared;
  k.work = a->work;
  k.peaks = a->peaks;
  k.peak_capacity = a->peak_capacity;
  k.peak_count = a->peak_count;
  return k;
}

struct Node {
  uint32_t start;
  uint32_t size;
};

// The node at `depth` below a node (start, size) that contains row r, or the
// leaf above it if the recursion stops first (size 1).
__device__ __forceinline__ Node find_node(uint32_t start, uint32_t size, uint32_t depth,
                                          uint32_t r) {
  for (uint32_t d = 0; d < depth && size > 1; ++d) {
    const uint32_t h = size / 2;
    if (r < start + h) {
      size = h;
    } else {
      start += h;
      size -= h;
    }
  }
  return Node{start, size};
}

__device__ __forceinline__ const float* task_input(const KArgs& a,
                                                   const gaffa_hip_ffa_group& g,
                                                   uint64_t series) {
  if (g.factor == 1.0) return a.series + series * a.series_stride;
  return a.prepared + series * a.plan.prepared_stride + g.prep_offset;
}

// ------------------------------------------------------------- prepare ----
__global__ void __launch_bounds__(256)
ffa_prepare_kernel(const float* __restrict__ series, uint64_t series_stride,
                   uint64_t nsamples, const gaffa_hip_ffa_group* __restrict__ groups,
                   float* __restrict__ prepared, uint64_t prepared_stride) {
  const gaffa_hip_ffa_group g = groups[blockIdx.y];
  if (g.factor == 1.0) return;
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (k >= g.prepared_nsamples) return;
  const uint64_t sidx = blockIdx.z;
  const float* x = series + sidx * series_stride;
  const double last_index = static_cast<double>(nsamples - 1);
  const double start = static_cast<double>(k) * g.factor;
  const double end = start + g.factor;
  const uint64_t first = static_cast<uint64_t>(floor(start));
  const uint64_t last = static_cast<uint64_t>(fmin(floor(end), last_index));
  const float w0 = static_cast<float>(static_cast<double>(first + 1) - start);
  const float w1 = static_cast<float>(end - static_cast<double>(last));
  float sum = w0 * x[first];
  for (uint64_t i = first + 1; i < last; ++i) sum += x[i];
  sum += w1 * x[last];
  prepared[sidx * prepared_stride + g.prep_offset + k] = sum;
}

// ------------------------------------------------------------- subtree ----
// One node at the task's LDS depth: load, run every level below it in LDS,
// store the node's transform into work[depth % 2].
__global__ void __launch_bounds__(kSubtreeThreads)
ffa_subtree_kernel(KArgs a) {
  extern __shared__ float lds[];
  const gaffa_hip_ffa_plan_desc& plan = a.plan;
  const gaffa_hip_ffa_item item = plan.subtree_items[blockIdx.x];
  const gaffa_hip_ffa_task task = plan.tasks[item.task];
  const gaffa_hip_ffa_group group = plan.groups[task.group];
  const uint64_t sidx = blockIdx.y;
  const uint32_t b = task.bins;
  const uint32_t nrows = item.b;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = tid >> 6;
  constexpr uint32_t kWaves = kSubtreeThreads / 64;

  float* buf0 = lds;
  float* buf1 = lds + static_cast<size_t>(nrows) * b;
  const float* in = task_input(a, group, sidx) + static_cast<uint64_t>(item.a) * b;
  for (uint32_t e = tid; e < nrows * b; e += kSubtreeThreads) buf0[e] = in[e];
  __syncthreads();

  uint32_t levels = 0;
  for (uint32_t sz = nrows; sz > 1; sz = sz - sz / 2) ++levels;  // ceil(log2(nrows))
  float* cur = buf0;
  float* nxt = buf1;
  for (int32_t d = static_cast<int32_t>(levels) - 1; d >= 0; --d) {
    for (uint32_t r = wave; r < nrows; r += kWaves) {
      const Node nd = find_node(0, nrows, static_cast<uint32_t>(d), r);
      float* o = nxt + static_cast<size_t>(r) * b;
      if (nd.size < 2) {
        for (uint32_t p = lane; p < b; p += 64) o[p] = cur[static_cast<size_t>(r) * b + p];
        continue;
      }
      const uint32_t s = r - nd.start;
      const uint32_t hn = nd.size / 2;
      const uint32_t tn = nd.size - hn;
      const uint32_t h = pick_row(hn, nd.size, s);
      const uint32_t t = pick_row(tn, nd.size, s);
      const uint32_t roll = (s - t) % b;
      const float* hr = cur + static_cast<size_t>(nd.start + h) * b;
      const float* tr = cur + static_cast<size_t>(nd.start + hn + t) * b;
      for (uint32_t p = lane; p < b; p += 64) {
        uint32_t j = p + roll;
        j = min(j, j - b);  // (p + roll) mod b for p, roll < b
        o[p] = hr[p] + tr[j];
      }
    }
    __syncthreads();
    float* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  float* out = a.work + static_cast<uint64_t>(task.depth & 1u) * a.nseries * plan.work_stride +
               sidx * plan.work_stride + task.work_offset + static_cast<uint64_t>(item.a) * b;
  for (uint32_t e = tid; e < nrows * b; e += kSubtreeThreads) out[e] = cur[e];
}

// --------------------------------------------------------------- level ----
// Rows [first, first + count) of the depth-`depth` nodes of a task, merged
// from the depth+1 outputs in work[(depth+1) % 2] into work[depth % 2].
__global__ void __launch_bounds__(kLevelThreads)
ffa_level_kernel(KArgs a, uint32_t item_offset) {
  const gaffa_hip_ffa_plan_desc& plan = a.plan;
  const gaffa_hip_ffa_item item = plan.level_items[item_offset + blockIdx.x];
  const gaffa_hip_ffa_task task = plan.tasks[item.task];
  const uint64_t sidx = blockIdx.y;
  const uint32_t b = task.bins;
  const uint32_t depth = item.a;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  const uint64_t region = a.nseries * plan.work_stride;
  const float* src = a.work + static_cast<uint64_t>((depth + 1) & 1u) * region +
                     sidx * plan.work_stride + task.work_offset;
  float* dst = a.work + static_cast<uint64_t>(depth & 1u) * region + sidx * plan.work_stride +
               task.work_offset;
  for (uint32_t i = wave; i < item.c; i += kLevelThreads / 64) {
    const uint32_t r = item.b + i;
    const Node nd = find_node(0, task.rows, depth, r);
    const uint32_t s = r - nd.start;
    const uint32_t hn = nd.size / 2;
    const uint32_t tn = nd.size - hn;
    const uint32_t h = pick_row(hn, nd.size, s);
    const uint32_t t = pick_row(tn, nd.size, s);
    const uint32_t roll = (s - t) % b;
    const float* hr = src + static_cast<uint64_t>(nd.start + h) * b;
    const float* tr = src + static_cast<uint64_t>(nd.start + hn + t) * b;
    float* o = dst + static_cast<uint64_t>(r) * b;
    for (uint32_t p = lane; p < b; p += 64) {
      uint32_t j = p + roll;
      j = min(j, j - b);
      o[p] = hr[p] + tr[j];
    }
  }
}

// ------------------------------------------------------------ terminal ----
template <bool kDetect>
__global__ void __launch_bounds__(kTerminalThreads)
ffa_terminal_kernel(KArgs a, float* __restrict__ rows_out,
                    uint32_t only_task) {
  extern __shared__ float lds[];
  const gaffa_hip_ffa_plan_desc& plan = a.plan;
  const gaffa_hip_ffa_item item = plan.terminal_items[blockIdx.x];
  if (!kDetect && item.task != only_task) return;
  const gaffa_hip_ffa_task task = plan.tasks[item.task];
  const gaffa_hip_ffa_group group = plan.groups[task.group];
  const uint64_t sidx = blockIdx.y;
  const uint32_t b = task.bins;
  const uint32_t s0 = item.a;
  const uint32_t nrows = item.b;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t wave = threadIdx.x >> 6;
  constexpr uint32_t kWaves = kTerminalThreads / 64;
  uint32_t maxw = 0;
  for (uint32_t w = 0; w < task.nwidths; ++w) maxw = max(maxw, task.width[w]);
  const uint32_t sb = (b + maxw) | 1u;  // odd stride: lane-per-row reads hit distinct banks

  // Root merge of rows [s0, s0 + nrows) into LDS, with circular extension.
  const float* src;
  uint32_t hn = 0, tn = 0;
  if (task.rows == 1) {
    src = task_input(a, group, sidx);
  } else {
    src = a.work + 1ull * a.nseries * plan.work_stride + sidx * plan.work_stride +
          task.work_offset;
    hn = task.rows / 2;
    tn = task.rows - hn;
  }
  for (uint32_t i = wave; i < nrows; i += kWaves) {
    const uint32_t s = s0 + i;
    float* o = lds + static_cast<size_t>(i) * sb;
    if (task.rows == 1) {
      for (uint32_t p = lane; p < b; p += 64) o[p] = src[p];
    } else {
      const uint32_t h = pick_row(hn, task.rows, s);
      const uint32_t t = pick_row(tn, task.rows, s);
      const uint32_t roll = (s - t) % b;
      const float* hr = src + static_cast<uint64_t>(h) * b;
      const float* tr = src + static_cast<uint64_t>(hn + t) * b;
      for (uint32_t p = lane; p < b; p += 64) {
        uint32_t j = p + roll;
        j = min(j, j - b);
        o[p] = hr[p] + tr[j];
      }
    }
  }
  __syncthreads();
  if constexpr (!kDetect) {
    for (uint32_t i = wave; i < nrows; i += kWaves)
      for (uint32_t p = lane; p < b; p += 64)
        rows_out[(sidx * task.rows_eval + s0 + i) * b + p] = lds[static_cast<size_t>(i) * sb + p];
    return;
  } else {
    for (uint32_t i = wave; i < nrows; i += kWaves)
      for (uint32_t p = lane; p < maxw; p += 64)
        lds[static_cast<size_t>(i) * sb + b + p] = lds[static_cast<size_t>(i) * sb + p];
    __syncthreads();

    // One lane per (row, width): wave-uniform width, lanes = rows.
    const uint32_t npairs = kTerminalRows * task.nwidths;
    for (uint32_t e = threadIdx.x; e < npairs; e += kTerminalThreads) {
      const uint32_t i = e % kTerminalRows;
      const uint32_t widx = e / kTerminalRows;
      if (i >= nrows) continue;
      const uint32_t w = task.width[widx];
      const float* x = lds + static_cast<size_t>(i) * sb;
      float A = 0.0f;  // P[phi]
      float B = 0.0f;  // P[phi + w]
      for (uint32_t k = 0; k < w; ++k) B += x[k];
      float best = -INFINITY;
      for (uint32_t p = 0; p < b; ++p) {
        best = fmaxf(best, B - A);
        A += x[p];
        B += x[p + w];
      }
      const float hb = task.height[widx] + task.baseline[widx];
      const float snr = (hb * best - task.baseline[widx] * A) * task.inv_stdnoise;
      if (snr >= a.snr_threshold) {
        // first phase attaining the maximum (ffa_detection.cpp:158-167)
        float A2 = 0.0f, B2 = 0.0f;
        for (uint32_t k = 0; k < w; ++k) B2 += x[k];
        uint32_t phase = 0;
        for (uint32_t p = 0; p < b; ++p) {
          if (B2 - A2 == best) {
            phase = p;
            break;
          }
          A2 += x[p];
          B2 += x[p + w];
        }
        const uint32_t slot = atomicAdd(a.peak_count, 1u);
        if (slot < a.peak_capacity) {
          gaffa_hip_ffa_peak pk;
          pk.series = a.series_index_base + static_cast<uint32_t>(sidx);
          pk.task = item.task;
          pk.shift = s0 + i;
          pk.phase_width = phase | (widx << 16);
          pk.snr = snr;
          a.peaks[slot] = pk;
        }
      }
    }
  }
}

void validate(const gaffa_hip_ffa_search_args* a) {
  check(a != nullptr && a->plan != nullptr, "FFA arguments must not be null");
  const gaffa_hip_ffa_plan_desc& p = *a->plan;
  check(p.groups != nullptr && p.tasks != nullptr && p.terminal_items != nullptr,
        "FFA plan tables must not be null");
  check(p.ntasks > 0, "FFA executor plan must contain at least one task");
  check(a->series != nullptr, "FFA series must not be null");
  check(a->nseries > 0 && a->nseries < 65536, "FFA batch must hold 1..65535 series");
  check(a->nsamples > 0, "FFA executor time series must not be empty");
  check(a->series_stride >= a->nsamples, "FFA series stride must be >= nsamples");
  check(a->work != nullptr, "FFA work buffer must not be null");
  check(p.max_bins > 1 && p.max_bins < 65536, "FFA bins must be in (1, 65536)");
}

}  // namespace
}  // namespace gaffa_hip

using namespace gaffa_hip;

namespace {

void launch_prepare(const gaffa_hip_ffa_search_args* a, hipStream_t st) {
  const gaffa_hip_ffa_plan_desc& p = *a->plan;
  if (p.ngroups == 0) return;
  // host-side maximum of prepared_nsamples is carried in prepared_stride
  const uint64_t maxp = p.max_prepared_nsamples;
  if (maxp == 0) return;
  check(a->prepared != nullptr, "FFA prepared buffer must not be null");
  const dim3 grid(static_cast<uint32_t>(ceil_div(maxp, 256)), p.ngroups,
                  static_cast<uint32_t>(a->nseries));
  hipLaunchKernelGGL(ffa_prepare_kernel, grid, dim3(256), 0, st, a->series, a->series_stride,
                     a->nsamples, p.groups, a->prepared, p.prepared_stride);
  launch_check("ffa_prepare_kernel");
}

void launch_transform(const gaffa_hip_ffa_search_args* a, hipStream_t st) {
  const gaffa_hip_ffa_plan_desc& p = *a->plan;
  const uint32_t ns = static_cast<uint32_t>(a->nseries);
  if (p.nsubtree_items > 0) {
    check(p.subtree_lds_bytes <= kMaxLdsBytes, "FFA subtree exceeds LDS");
    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(ffa_subtree_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(p.subtree_lds_bytes)),
              "hipFuncSetAttribute");
    hipLaunchKernelGGL(ffa_subtree_kernel, dim3(p.nsubtree_items, ns), dim3(kSubtreeThreads),
                       p.subtree_lds_bytes, st, kargs(a));
    launch_check("ffa_subtree_kernel");
  }
  for (int32_t d = static_cast<int32_t>(p.max_depth) - 1; d >= 1; --d) {
    const uint32_t lo = p.level_item_offsets[d];
    const uint32_t hi = p.level_item_offsets[d + 1];
    if (hi <= lo) continue;
    hipLaunchKernelGGL(ffa_level_kernel, dim3(hi - lo, ns), dim3(kLevelThreads), 0, st, kargs(a),
                       lo);
    launch_check("ffa_level_kernel");
  }
}

uint32_t terminal_lds_bytes(const gaffa_hip_ffa_plan_desc& p) {
  // rows of (max_bins + max width) floats, odd stride; widths < bins
  const uint32_t sb = (2 * p.max_bins) | 1u;
  return kTerminalRows * sb * 4;
}

}  // namespace

extern "C" {

int gaffa_hip_ffa_prepare_batch(const gaffa_hip_ffa_search_args* a, gaffa_hip_stream stream) {
  return guarded([&] {
    validate(a);
    launch_prepare(a, as_stream(stream));
  });
}

int gaffa_hip_ffa_search_batch(const gaffa_hip_ffa_search_args* a, gaffa_hip_stream stream) {
  return guarded([&] {
    validate(a);
    check(a->peaks != nullptr && a->peak_count != nullptr, "FFA peak buffers must not be null");
    check(std::isfinite(a->snr_threshold), "FFA search S/N threshold must be finite");
    const hipStream_t st = as_stream(stream);
    launch_prepare(a, st);
    launch_transform(a, st);
    const gaffa_hip_ffa_plan_desc& p = *a->plan;
    const uint32_t lds = terminal_lds_bytes(p);
    check(lds <= kMaxLdsBytes, "FFA bins too large for the detection tile");
    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(ffa_terminal_kernel<true>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds)),
              "hipFuncSetAttribute");
    hipLaunchKernelGGL(ffa_terminal_kernel<true>,
                       dim3(p.nterminal_items, static_cast<uint32_t>(a->nseries)),
                       dim3(kTerminalThreads), lds, st, kargs(a), static_cast<float*>(nullptr),
                       0u);
    launch_check("ffa_terminal_kernel");
  });
}

int gaffa_hip_ffa_transform_rows(const gaffa_hip_ffa_search_args* a, uint32_t task, float* out,
                                 gaffa_hip_stream stream) {
  return guarded([&] {
    validate(a);
    check(out != nullptr, "FFA transform output must not be null");
    check(task < a->plan->ntasks, "FFA task index out of range");
    const hipStream_t st = as_stream(stream);
    launch_prepare(a, st);
    launch_transform(a, st);
    const gaffa_hip_ffa_plan_desc& p = *a->plan;
    const uint32_t lds = terminal_lds_bytes(p);
    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(ffa_terminal_kernel<false>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds)),
              "hipFuncSetAttribute");
    hipLaunchKernelGGL(ffa_terminal_kernel<false>,
                       dim3(p.nterminal_items, static_cast<uint32_t>(a->nseries)),
                       dim3(kTerminalThreads), lds, st, kargs(a), out, task);
    launch_check("ffa_terminal_kernel");
  });
}

}  // extern "C"
