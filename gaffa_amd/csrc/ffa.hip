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
//   prepare   riptide weighted-sum downsample (time_series.cpp:140-174): a
//             workgroup stages a contiguous input span in LDS with coalesced
//             loads, then each output is summed serially in the CPU's FP32
//             order
//   subtree   persistent, one workgroup per CU walking (node at the task's
//             LDS depth D, series) units: the node's rows live in LDS and
//             every level below it runs there, ping-ponging between two LDS
//             buffers, with merges from host-built tables
//   level     one launch per output depth: one to three levels per pass,
//             each item (32 rows of one node) copying its leaf rows' runs
//             from HBM into LDS and merging there
//   terminal  one workgroup per (32 final rows, series): the root merge, or
//             the top two levels where that saves a global pass, into LDS
//             (rows stored with a circular extension), then one lane per
//             (row, width) runs the serial FP32 prefix twice, P[phi] and
//             P[phi + w], so every boxcar sum equals the CPU's circular-prefix
//             difference bit for bit (ffa_detection.cpp:126-239); significant
//             peaks are appended with one atomic each.
// The subtree and level kernels read their merges and leaves from host-built
// tables; in the terminal kernel the (head row, tail row, roll) descriptor of
// each output row is computed once, by one lane, into LDS.  The waves then
// stream whole rows (lanes over phase bins) reading only those.  The subtree
// and level passes write their rows with streaming stores.
#include "common.hpp"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace gaffa_hip {
namespace {

constexpr uint32_t kLevelThreads = 512;
constexpr uint32_t kLevelRows = 32;     // default rows per level work item
constexpr uint32_t kLevelRowsMax = 32;  // most rows a level item may hold
constexpr uint32_t kSubtreeThreads = 1024;
constexpr uint32_t kSubtreeMaxRows = 512;  // rows of a subtree node
constexpr uint32_t kTerminalThreads = 256;
constexpr uint32_t kTerminalRows = 32;     // default rows per terminal work item
constexpr uint32_t kTerminalRowsMax = 32;  // most rows a terminal item may hold
constexpr uint32_t kPrepThreads = 256;
constexpr uint32_t kPrepStage = 8192;  // floats of input staged per prepare block
constexpr uint32_t kMaxLdsBytes = 160 * 1024;

// ffa.cpp:133-143 in FP32, no contraction (the library is built with
// -ffp-contract=off).
__host__ __device__ __forceinline__ uint32_t pick_row(uint32_t child_rows, uint32_t rows,
                                             uint32_t s) {
  const float scale = static_cast<float>(child_rows - 1) / static_cast<float>(rows - 1);
  return static_cast<uint32_t>(scale * static_cast<float>(s) + 0.5f);
}

// base[idx] with the byte offset computed in 32 bits (idx < 2^30: a task's
// rows * bins, which the host checks), so the load is the uniform base in
// scalar registers plus one 32-bit lane offset.
__device__ __forceinline__ float ld32(const float* base, uint32_t idx) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + (idx << 2));
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
  float* periodogram;
  uint64_t periodogram_stride;
  uint32_t periodogram_widths;
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
  k.periodogram = a->periodogram;
  k.periodogram_stride = a->periodogram_stride;
  k.periodogram_widths = a->periodogram_widths;
  return k;
}

struct Node {
  uint32_t start;
  uint32_t size;
};

// The node at `depth` below a node (start, size) that contains row r, or the
// leaf above it if the recursion stops first (size 1).
__host__ __device__ __forceinline__ Node find_node(uint32_t start, uint32_t size,
                                                   uint32_t depth, uint32_t r) {
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

// Merge descriptor of row r of node (start, size >= 2): the head and tail
// child rows it reads and the roll applied to the tail row.
struct Desc {
  uint32_t h;
  uint32_t t;
  uint32_t roll;
};

__host__ __device__ __forceinline__ Desc merge_desc(Node nd, uint32_t r, uint32_t bins) {
  const uint32_t s = r - nd.start;
  const uint32_t hn = nd.size / 2;
  const uint32_t tn = nd.size - hn;
  const uint32_t h = pick_row(hn, nd.size, s);
  const uint32_t t = pick_row(tn, nd.size, s);
  return Desc{nd.start + h, nd.start + hn + t, (s - t) % bins};
}

__device__ __forceinline__ const float* task_input(const KArgs& a, const gaffa_hip_ffa_group& g,
                                                   uint64_t series) {
  if (g.factor == 1.0) return a.series + series * a.series_stride + g.input_offset;
  return a.prepared + series * a.plan.prepared_stride + g.prep_offset;
}

// One wave merges RB rows i = first + k * step (k < RB; rows >= count are
// skipped): out_i[p] = h_i[p] + t_i[(p + roll_i) mod b], lanes over p.  Every
// load address is clamped in bounds and issued unconditionally, so all
// 2 * RB * 4 loads of a batch are in flight together; only the stores are
// predicated.  `row(i, h, t, roll)` supplies row i's operands as element
// offsets from `base` (< 2^30) and is only called with i < count (skipped
// rows reuse row `first`'s operands).
template <uint32_t RB, typename RowFn, typename OutFn>
__device__ __forceinline__ void merge_batch(const float* base, uint32_t first, uint32_t step,
                                            uint32_t count, uint32_t b, uint32_t lane,
                                            RowFn&& row, OutFn&& out) {
  if (first >= count) return;  // wave-uniform: no row of this batch exists
  uint32_t hp[RB];
  uint32_t tp[RB];
  uint32_t roll[RB];
#pragma unroll
  for (uint32_t k = 0; k < RB; ++k) {
    const uint32_t i = first + k * step;
    row(i < count ? i : first, hp[k], tp[k], roll[k]);
  }
  for (uint32_t p0 = lane; p0 < b; p0 += 256) {
    float v[RB][4];
#pragma unroll
    for (uint32_t k = 0; k < RB; ++k) {
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t p = min(p0 + 64u * j, b - 1);
        uint32_t q = p + roll[k];
        q = min(q, q - b);  // (p + roll) mod b for p, roll < b
        v[k][j] = ld32(base, hp[k] + p) + ld32(base, tp[k] + q);
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < RB; ++k) {
      const uint32_t i = first + k * step;
#pragma unroll
      for (uint32_t j = 0; j < 4; ++j) {
        const uint32_t p = p0 + 64u * j;
        if (i < count && p < b) out(i, p, v[k][j]);
      }
    }
  }
}

// dst[e] = src[e] for e < total by `threads` threads, U loads in flight per
// thread (addresses clamped, stores predicated).
template <uint32_t U, typename D, typename S>
__device__ __forceinline__ void copy_span(D* dst, const S* src, uint32_t total, uint32_t tid,
                                          uint32_t threads) {
  for (uint32_t e0 = 0; e0 < total; e0 += U * threads) {
    float v[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) v[u] = src[min(e0 + u * threads + tid, total - 1)];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t e = e0 + u * threads + tid;
      if (e < total) dst[e] = v[u];
    }
  }
}

// ------------------------------------------------------------- prepare ----
// Outputs per block: K = min(256, 8192 / factor) so the staged input span
// (about K * factor samples) fits kPrepStage floats.
__device__ __host__ __forceinline__ uint32_t prep_outputs_per_block(double factor) {
  const double k = static_cast<double>(kPrepStage - 8) / factor;
  return k >= 256.0 ? 256u : (k < 1.0 ? 1u : static_cast<uint32_t>(k));
}

__global__ void __launch_bounds__(kPrepThreads)
ffa_prepare_kernel(const float* __restrict__ series, uint64_t series_stride,
                   const gaffa_hip_ffa_group* __restrict__ groups,
                   float* __restrict__ prepared, uint64_t prepared_stride) {
  __shared__ float stage[kPrepStage];
  const gaffa_hip_ffa_group g = groups[blockIdx.y];
  if (g.factor == 1.0) return;
  const uint32_t K = prep_outputs_per_block(g.factor);
  const uint64_t k0 = static_cast<uint64_t>(blockIdx.x) * K;
  if (k0 >= g.prepared_nsamples) return;
  const uint64_t k1 = min(k0 + K, g.prepared_nsamples);
  const uint64_t sidx = blockIdx.z;
  const float* x = series + sidx * series_stride + g.input_offset;
  const double last_index = static_cast<double>(g.input_nsamples - 1);
  // staged span: [floor(k0 f), min(floor(k1 f) + 2, n - 1)]
  const uint64_t lo = static_cast<uint64_t>(floor(static_cast<double>(k0) * g.factor));
  const uint64_t hi = static_cast<uint64_t>(
      fmin(floor(static_cast<double>(k1) * g.factor) + 2.0, last_index));
  // K * factor <= kPrepStage - 8 whenever K > 1; a single huge-factor output
  // (factor > kPrepStage - 8) reads global memory directly.
  const uint64_t span = hi - lo + 1;
  const bool staged = span <= kPrepStage;
  if (staged) {
    for (uint32_t e = threadIdx.x; e < span; e += kPrepThreads) stage[e] = x[lo + e];
    __syncthreads();
  }
  const uint64_t k = k0 + threadIdx.x;
  if (k >= k1) return;
  const double start = static_cast<double>(k) * g.factor;
  const double end = start + g.factor;
  const uint64_t first = static_cast<uint64_t>(floor(start));
  const uint64_t last = static_cast<uint64_t>(fmin(floor(end), last_index));
  const float w0 = static_cast<float>(static_cast<double>(first + 1) - start);
  const float w1 = static_cast<float>(end - static_cast<double>(last));
  // the serial sum, written once per source so each keeps its address space
  auto weighted = [&](const float* src, uint64_t base) {
    float sum = w0 * src[first - base];
    for (uint64_t i = first + 1; i < last; ++i) sum += src[i - base];
    return sum + w1 * src[last - base];
  };
  const float sum = staged ? weighted(stage, lo) : weighted(x, 0);
  prepared[sidx * prepared_stride + g.prep_offset + k] = sum;
}

// All downsample groups of a series in one pass over it: block (span x,
// series z) stages input samples [s0, s0 + kPrepSpan + halo) once and
// computes every output k of every group whose window starts in
// [s0, s0 + kPrepSpan) (floor(k f) in the span; the window reaches at most
// ceil(f) + 1 samples further, inside the halo), with the same serial
// weighted sum as ffa_prepare_kernel.  For plans whose downsample groups all
// read one series (input_offset 0, one input length) and whose largest
// factor fits the halo; the series is then read once instead of once per
// group.
constexpr uint32_t kPrepSpan = 8192;
constexpr uint32_t kPrepMaxHalo = 4096;
constexpr uint32_t kPrepMaxGroups = 64;

__global__ void __launch_bounds__(kPrepThreads)
ffa_prepare_fused_kernel(const float* __restrict__ series, uint64_t series_stride,
                         const gaffa_hip_ffa_group* __restrict__ groups, uint32_t ngroups,
                         uint32_t halo, float* __restrict__ prepared, uint64_t prepared_stride) {
  extern __shared__ float stage[];
  __shared__ uint64_t k_lo[kPrepMaxGroups];
  __shared__ uint32_t first_out[kPrepMaxGroups + 1];
  const uint64_t n = groups[0].input_nsamples;
  const uint64_t s0 = static_cast<uint64_t>(blockIdx.x) * kPrepSpan;
  if (s0 >= n) return;
  const uint64_t s1 = min(s0 + kPrepSpan, n);  // owned window starts
  const uint64_t sidx = blockIdx.z;
  const float* x = series + sidx * series_stride;
  const uint64_t staged = min(n, s1 + halo) - s0;
  for (uint32_t e = threadIdx.x; e < staged; e += kPrepThreads) stage[e] = x[s0 + e];
  // per group: the outputs whose window starts in [s0, s1)
  if (threadIdx.x < ngroups) {
    const gaffa_hip_ffa_group g = groups[threadIdx.x];
    uint64_t lo = 0, hi = 0;
    if (g.factor != 1.0) {
      const double f = g.factor;
      auto start = [&](uint64_t k) { return static_cast<uint64_t>(floor(static_cast<double>(k) * f)); };
      lo = static_cast<uint64_t>(ceil(static_cast<double>(s0) / f));
      while (lo > 0 && start(lo - 1) >= s0) --lo;
      while (start(lo) < s0) ++lo;
      hi = static_cast<uint64_t>(ceil(static_cast<double>(s1) / f));
      while (hi > 0 && start(hi - 1) >= s1) --hi;
      while (start(hi) < s1) ++hi;
      lo = min(lo, g.prepared_nsamples);
      hi = min(hi, g.prepared_nsamples);
    }
    k_lo[threadIdx.x] = lo;
    first_out[threadIdx.x + 1] = static_cast<uint32_t>(hi - lo);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    first_out[0] = 0;
    for (uint32_t gi = 0; gi < ngroups; ++gi) first_out[gi + 1] += first_out[gi];
  }
  __syncthreads();
  const double last_index = static_cast<double>(n - 1);
  const uint32_t total = first_out[ngroups];
  uint32_t gi = 0;
  for (uint32_t e = threadIdx.x; e < total; e += kPrepThreads) {
    while (first_out[gi + 1] <= e) ++gi;
    const gaffa_hip_ffa_group g = groups[gi];
    const uint64_t k = k_lo[gi] + (e - first_out[gi]);
    const double start = static_cast<double>(k) * g.factor;
    const double end = start + g.factor;
    const uint64_t first = static_cast<uint64_t>(floor(start));
    const uint64_t last = static_cast<uint64_t>(fmin(floor(end), last_index));
    const float w0 = static_cast<float>(static_cast<double>(first + 1) - start);
    const float w1 = static_cast<float>(end - static_cast<double>(last));
    const float* src = stage - s0;
    float sum = w0 * src[first];
    for (uint64_t i = first + 1; i < last; ++i) sum += src[i];
    sum += w1 * src[last];
    prepared[sidx * prepared_stride + g.prep_offset + k] = sum;
  }
}

// ------------------------------------------------------------- subtree ----
// Persistent: one workgroup per CU walks the (node, series) units of the
// launch, unit u = (subtree item u % nitems, series u / nitems), so that the
// units in flight together share a series' prepared data in L2.  For each
// unit: the node's rows, prefetched into registers while the previous unit
// computed, are written to LDS; the levels run in LDS (ping-pong), the top
// level writing straight to the task's work region.  The merges come from
// the item's host-built table (gaffa_hip_ffa_merge), read with scalar loads
// one pass ahead.  LDS = zero rows (subtree_zero_floats(max_bins) floats of
// -0.0), [2][half] floats of rows, kSubtreeGuard floats of slack.  The zero
// rows are the tail of a copied row (x + -0.0f == x for every x) and keep
// the merge's tail base addresses positive; the slack holds the last-chunk
// overreads.
constexpr uint32_t kSubtreePrefetch = 20;  // floats per thread: node <= 20480 floats
constexpr uint32_t kSubtreeGuard = 256;    // floats of slack after the row buffers
constexpr uint32_t kSubtreeRowsPerPass = 2;  // rows a wave merges at once (table layout too)
constexpr uint32_t kSubtreeWaves = kSubtreeThreads / 64;

__host__ __device__ __forceinline__ uint32_t tree_levels(uint32_t nrows) {
  uint32_t levels = 0;
  for (uint32_t sz = nrows; sz > 1; sz = sz - sz / 2) ++levels;  // ceil(log2(nrows))
  return levels;
}

// Floats of -0.0 in front of the row buffers: a copied row's tail reads
// phases up to bins + 62, and a tail base may sit 192 floats before its row.
__host__ __device__ __forceinline__ uint32_t subtree_zero_floats(uint32_t max_bins) {
  const uint32_t need = max_bins + 64u > 256u ? max_bins + 64u : 256u;
  return (need + 63u) & ~63u;
}

// Passes over a subtree node of `levels` tree levels: two levels each, the
// bottom one single when `levels` is odd.
__host__ __device__ __forceinline__ uint32_t subtree_tree_passes(uint32_t levels) {
  return (levels + 1) / 2;
}

// Passes per wave per level of a node of `rows` rows: wave w merges rows w,
// w + 16, ..., kSubtreeRowsPerPass of them per pass.
__host__ __device__ __forceinline__ uint32_t subtree_passes(uint32_t rows) {
  const uint32_t per_wave = (rows + kSubtreeWaves - 1) / kSubtreeWaves;
  return (per_wave + kSubtreeRowsPerPass - 1) / kSubtreeRowsPerPass;
}

using ConstMerge = __attribute__((address_space(4))) const gaffa_hip_ffa_merge;

// A merge read through the constant address space: a uniform address, so
// scalar loads.
__device__ __forceinline__ gaffa_hip_ffa_merge load_merge(ConstMerge* p) {
  gaffa_hip_ffa_merge g;
  g.head = p->head;
  g.tail = p->tail;
  g.wrap = p->wrap;
  g.out = p->out;
  g.head2 = p->head2;
  g.wrap2 = p->wrap2;
  g.tail2 = p->tail2;
  g.wrap3 = p->wrap3;
  return g;
}

__global__ void __launch_bounds__(kSubtreeThreads)
ffa_subtree_kernel(KArgs a, uint32_t nunits) {
  extern __shared__ float lds[];
  const gaffa_hip_ffa_plan_desc& plan = a.plan;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  constexpr uint32_t kWaves = kSubtreeWaves;
  const uint32_t half = plan.subtree_half_floats;
  const uint32_t zero_floats = subtree_zero_floats(plan.max_bins);
  float* const rows = lds + zero_floats;
  const uint32_t nitems = plan.nsubtree_items;
  ConstMerge* const merges = (ConstMerge*)plan.subtree_merges;
  // The plan tables are read through the constant address space (scalar
  // loads), so that the only vector loads in flight are the prefetched rows
  // and no wait for plan fields drains them early.
  using ConstItem = __attribute__((address_space(4))) const gaffa_hip_ffa_item;
  using ConstTask = __attribute__((address_space(4))) const gaffa_hip_ffa_task;
  using ConstGroup = __attribute__((address_space(4))) const gaffa_hip_ffa_group;
  ConstItem* const items = (ConstItem*)plan.subtree_items;
  ConstTask* const tasks = (ConstTask*)plan.tasks;
  ConstGroup* const groups = (ConstGroup*)plan.groups;

  for (uint32_t e = tid; e < zero_floats; e += kSubtreeThreads) lds[e] = -0.0f;

  float pf[kSubtreePrefetch];
  auto fetch = [&](uint32_t u) {
    ConstItem* it = items + u % nitems;
    ConstTask* t = tasks + it->task;
    ConstGroup* g = groups + t->group;
    const uint64_t series = u / nitems;
    const float* in = (g->factor == 1.0
                           ? a.series + series * a.series_stride + g->input_offset
                           : a.prepared + series * plan.prepared_stride + g->prep_offset) +
                      static_cast<uint64_t>(it->a) * t->bins;
    const uint32_t total = it->b * t->bins;  // >= 1
#pragma unroll
    for (uint32_t k = 0; k < kSubtreePrefetch; ++k)
      pf[k] = in[min(tid + k * kSubtreeThreads, total - 1)];
  };

  uint32_t u = blockIdx.x;
  if (u >= nunits) return;
  fetch(u);
  for (; u < nunits; u += gridDim.x) {
    ConstItem* const itp = items + u % nitems;
    ConstTask* const tp = tasks + itp->task;
    const uint32_t node_start = itp->a, nrows = itp->b, first_merge = itp->c;
    const uint32_t b = tp->bins;
    const uint32_t region = (tp->regions >> tp->depth) & 1u;
    const uint64_t work_offset = tp->work_offset;
    const uint64_t sidx = u / nitems;
    const uint32_t total = nrows * b;
    float* cur = rows;
    float* nxt = rows + half;
#pragma unroll
    for (uint32_t k = 0; k < kSubtreePrefetch; ++k) {
      const uint32_t e = tid + k * kSubtreeThreads;
      if (e < total) cur[e] = pf[k];
    }
    const uint32_t levels = tree_levels(nrows);
    __syncthreads();
    // the next unit's rows travel while this one computes
    if (u + gridDim.x < nunits) fetch(u + gridDim.x);

    float* out = a.work + static_cast<uint64_t>(region) * a.nseries * plan.work_stride +
                 sidx * plan.work_stride + work_offset + static_cast<uint64_t>(node_start) * b;
    if (levels == 0) {
      for (uint32_t e = tid; e < total; e += kSubtreeThreads) out[e] = cur[e];
    }
    // One pass of NL = 1 or 2 tree levels: the wave's merges m[0 .. R *
    // passes) of `src` (LDS) into `store`, which is written once per
    // destination kind so every access keeps its address space (LDS rows are
    // ds_* operations, the top pass's stores global_*).  The table gives each
    // wave its rows w, w + 16, ..., padded to whole passes with entries that
    // are skipped.  A row is swept in blocks of 256 phases, NC <= 4 chunks of
    // 64 lanes.  A rolled operand X (B, and C, D in a two-level pass) splits
    // the row at its wrap point w = bins - roll: phases p < w read X[p +
    // roll], phases p >= w read X[p + roll - b].  Per chunk a lane picks one
    // of two addresses (one compare, one select); every other address is one
    // add per row plus the chunk's immediate offset.  The last chunk of a row
    // loads up to 63 floats past it (inside the next row or the LDS slack)
    // and masks only its stores.  A copied operand reads the zero rows.
    const uint32_t passes = subtree_passes(nrows);
    auto level = [&](auto nl, ConstMerge* m, const float* src, auto&& store) {
      constexpr uint32_t NL = decltype(nl)::value;
      if (wave >= nrows) return;
      constexpr uint32_t R = kSubtreeRowsPerPass;
      for (uint32_t p0 = 0; p0 < b; p0 += 256) {
        const uint32_t rem = b - p0;
        const uint32_t nc = rem >= 256 ? 4u : (rem + 63) >> 6;
        const uint32_t live = min(rem - 64u * (nc - 1), 64u);  // lanes of the last chunk
        auto pass = [&](auto ncc) {
          constexpr uint32_t NC = decltype(ncc)::value;
          using LdsFloat = __attribute__((address_space(3))) const float;
          // X[(p + roll) mod b] for the chunk's phases p0 + lane + 64 j.
          // tb[64 j] lies inside X's row for every lane that reads it
          // (p >= w); tb itself may sit up to 192 floats before the row,
          // inside the zero rows.  The empty asm keeps ta and tb two
          // registers, so that each chunk's choice is one select rather than
          // a rebuilt address.
          auto load_rolled = [&](uint32_t x, uint32_t wrap, float (&v)[NC]) {
            const uint32_t before = wrap - min(wrap, p0);  // phases of the block before w
            const float* base = x == GAFFA_HIP_FFA_COPY_ROW ? lds : src + x;
            LdsFloat* ta = (LdsFloat*)(base + p0 + lane);
            LdsFloat* tb = ta - b;
            asm("" : "+v"(tb));
#pragma unroll
            for (uint32_t j = 0; j < NC; ++j) {
              LdsFloat* tp = lane + 64u * j < before ? ta : tb;
              v[j] = tp[64u * j];
            }
          };
          auto load_row = [&](const gaffa_hip_ffa_merge& g, float (&v)[2 * NL][NC]) {
            const float* hp = src + g.head + p0 + lane;
#pragma unroll
            for (uint32_t j = 0; j < NC; ++j) v[0][j] = hp[64u * j];
            load_rolled(g.tail, g.wrap, v[1]);
            if constexpr (NL == 2) {
              load_rolled(g.head2, g.wrap2, v[2]);
              load_rolled(g.tail2, g.wrap3, v[3]);
            }
          };
          auto store_row = [&](uint32_t orow, const float (&v)[2 * NL][NC]) {
            auto sum = [&](uint32_t j) {
              if constexpr (NL == 2) return (v[0][j] + v[1][j]) + (v[2][j] + v[3][j]);
              else return v[0][j] + v[1][j];
            };
#pragma unroll
            for (uint32_t j = 0; j + 1 < NC; ++j) store(orow + p0 + lane + 64u * j, sum(j));
            if (lane < live) store(orow + p0 + lane + 64u * (NC - 1), sum(NC - 1));
          };
          // every load of the pass is issued before the first add waits; a
          // wave with fewer rows than the longest list skips its padding (the
          // last one or two entries of its list)
          static_assert(R == 2, "the padding cases below assume two rows per pass");
          auto run = [&](const gaffa_hip_ffa_merge (&cg)[R]) {
            if (cg[1].out != GAFFA_HIP_FFA_SKIP_ROW) {
              float v[R][2 * NL][NC];
#pragma unroll
              for (uint32_t k = 0; k < R; ++k) load_row(cg[k], v[k]);
              __builtin_amdgcn_sched_barrier(0);
#pragma unroll
              for (uint32_t k = 0; k < R; ++k) store_row(cg[k].out, v[k]);
            } else if (cg[0].out != GAFFA_HIP_FFA_SKIP_ROW) {
              float v[2 * NL][NC];
              load_row(cg[0], v);
              __builtin_amdgcn_sched_barrier(0);
              store_row(cg[0].out, v);
            }
          };
          // the merges of the next pass load while this one runs, into the
          // other of two register sets (unrolled by two, so that no set is
          // ever copied into the other)
          gaffa_hip_ffa_merge ga[R], gb[R];
#pragma unroll
          for (uint32_t k = 0; k < R; ++k) ga[k] = load_merge(m + k);
          for (uint32_t ps = 0; ps < passes; ps += 2) {
            const bool second = ps + 1 < passes;
            if (second) {
#pragma unroll
              for (uint32_t k = 0; k < R; ++k) gb[k] = load_merge(m + (ps + 1) * R + k);
            }
            run(ga);
            if (!second) break;
            if (ps + 2 < passes) {
#pragma unroll
              for (uint32_t k = 0; k < R; ++k) ga[k] = load_merge(m + (ps + 2) * R + k);
            }
            run(gb);
          }
        };
        if (nc == 4) pass(std::integral_constant<uint32_t, 4>{});
        else if (nc == 3) pass(std::integral_constant<uint32_t, 3>{});
        else if (nc == 2) pass(std::integral_constant<uint32_t, 2>{});
        else pass(std::integral_constant<uint32_t, 1>{});
      }
    };
    // tree pass k of the table (k = 0 the node's own): kWaves x (R * passes)
    // merges, wave-major; two levels per pass, the bottom pass single when
    // the level count is odd
    ConstMerge* const table = merges + first_merge + wave * kSubtreeRowsPerPass * passes;
    const uint32_t pass_size = kWaves * kSubtreeRowsPerPass * passes;
    const uint32_t npasses = subtree_tree_passes(levels);
    using One = std::integral_constant<uint32_t, 1>;
    using Two = std::integral_constant<uint32_t, 2>;
    for (int32_t k = static_cast<int32_t>(npasses) - 1; k >= 1; --k) {
      float* dst = nxt;
      auto to_lds = [&](uint32_t i, float v) { dst[i] = v; };
      ConstMerge* const mk = table + static_cast<uint32_t>(k) * pass_size;
      if (2u * static_cast<uint32_t>(k) + 1u < levels) level(Two{}, mk, cur, to_lds);
      else level(One{}, mk, cur, to_lds);
      __syncthreads();
      nxt = cur;
      cur = dst;
    }
    if (levels > 0) {
      // streaming stores: the node's rows are read back by the next kernel,
      // long after L2 would have had to hold them
      auto to_global = [&](uint32_t i, float v) { __builtin_nontemporal_store(v, &out[i]); };
      if (levels >= 2) level(Two{}, table, cur, to_global);
      else level(One{}, table, cur, to_global);
    }
    __syncthreads();  // the LDS rows are rewritten by the next unit
  }
}

// --------------------------------------------------------------- level ----

// One-level merge operands of a depth-d row for the two-level detection
// merge: (A[p] + B[p + rh]) + (C[q] + D[q + rt]) with q = p + ro (indices
// mod b), which is exactly the two one-level merges (the depth d + 1 values
// are FP32 sums in between), without writing the depth d + 1 rows.
struct Desc2 {
  uint32_t a, b, rh;  // head child row: its two inputs and roll
  uint32_t c, d, rt;  // tail child row: its two inputs and roll
  uint32_t ro;        // roll of the tail child row in the output merge
};

// A level pass of L = 1, 2 or 3 tree levels merges a depth-d row r straight
// from the 2^L depth-(d + L) rows on its paths down (its leaves), each at the
// combined roll R_j of its path (the sum of the rolls along it, mod b), in
// the order of the L one-level merges:
//   L = 1: v0 + v1
//   L = 2: (v0 + v1) + (v2 + v3)
//   L = 3: ((v0 + v1) + (v2 + v3)) + ((v4 + v5) + (v6 + v7))
// with v_j = leaf_j[(p + R_j) mod b]; the depth d + 1 .. d + L - 1 values are
// the FP32 sums in between and are never written.  The host builds each
// row's leaves (gaffa_hip_ffa_level_row) once per plan.
__host__ __device__ __forceinline__ gaffa_hip_ffa_level_row level_leaves(uint32_t task_rows,
                                                                         uint32_t depth,
                                                                         uint32_t levels,
                                                                         uint32_t r, uint32_t b) {
  gaffa_hip_ffa_level_row e{};
  const Node nd = find_node(0, task_rows, depth, r);
  const Desc top = merge_desc(nd, r, b);
  if (levels == 1) {
    e.row[0] = top.h;
    e.row[1] = top.t;
    e.roll[1] = top.roll;
    return e;
  }
  const uint32_t hn = nd.size / 2;
  const Node nh{nd.start, hn}, nt{nd.start + hn, nd.size - hn};
  const Desc dh = merge_desc(nh, top.h, b);
  const Desc dt = merge_desc(nt, top.t, b);
  if (levels == 2) {
    e.row[0] = dh.h;
    e.row[1] = dh.t;
    e.row[2] = dt.h;
    e.row[3] = dt.t;
    e.roll[1] = dh.roll;
    e.roll[2] = top.roll;
    e.roll[3] = (top.roll + dt.roll) % b;
    return e;
  }
  const Desc da = merge_desc(Node{nh.start, nh.size / 2}, dh.h, b);
  const Desc db = merge_desc(Node{nh.start + nh.size / 2, nh.size - nh.size / 2}, dh.t, b);
  const Desc dc = merge_desc(Node{nt.start, nt.size / 2}, dt.h, b);
  const Desc dd = merge_desc(Node{nt.start + nt.size / 2, nt.size - nt.size / 2}, dt.t, b);
  e.row[0] = da.h;
  e.row[1] = da.t;
  e.row[2] = db.h;
  e.row[3] = db.t;
  e.row[4] = dc.h;
  e.row[5] = dc.t;
  e.row[6] = dd.h;
  e.row[7] = dd.t;
  e.roll[1] = da.roll;
  e.roll[2] = dh.roll;
  e.roll[3] = (dh.roll + db.roll) % b;
  e.roll[4] = top.roll;
  e.roll[5] = (top.roll + dc.roll) % b;
  e.roll[6] = (top.roll + dt.roll) % b;
  e.roll[7] = (top.roll + dt.roll + dd.roll) % b;
  return e;
}

// Floats of slack after a level item's staged rows: the last chunk of a row
// reads up to 255 floats past it (lanes beyond the row, whose stores are
// masked).
constexpr uint32_t kLevelGuard = 256;
// Dynamic LDS a level launch may take: the LDS less the kernel's static
// leaf rows (kLevelRowsMax x 64 bytes).
constexpr uint32_t kMaxLevelLdsBytes = kMaxLdsBytes - 4096;

// Dynamic LDS an item of `count` rows from `first` needs to merge from LDS:
// for each leaf j, the run of depth-(d + L) rows from the smallest to the
// largest leaf_j over the item's rows (one short run per leaf when the item
// lies in one node), all runs back to back, plus the guard.
__host__ __device__ __forceinline__ uint64_t level_stage_bytes(uint32_t staged_rows, uint32_t b) {
  return (static_cast<uint64_t>(staged_rows) * b + kLevelGuard) * 4u;
}

// Rows [item.b, item.b + item.c) of the depth-d nodes of a task (item.a = d
// | L << 16).  The item's leaf runs (its span, host-built) are copied into
// LDS once (global_load_lds: 64 consecutive floats per instruction, straight
// into LDS) together with its rows' leaves (host-built), and each output
// phase then reads its 2^L operands there; a wave merges one row at a time,
// four 64-phase chunks per block.  An item whose runs exceed the launch's LDS
// (plan.level_lds_bytes) reads its operands from global memory instead, with
// the same sums.  The work region of each depth comes from the task's region
// bits.
__global__ void __launch_bounds__(kLevelThreads)
ffa_level_kernel(KArgs a, uint32_t item_offset) {
  extern __shared__ float stage[];
  using ConstSpan = __attribute__((address_space(4))) const gaffa_hip_ffa_level_span;
  const gaffa_hip_ffa_plan_desc& plan = a.plan;
  const uint32_t it = item_offset + blockIdx.x;
  const gaffa_hip_ffa_item item = plan.level_items[it];
  ConstSpan* const span = (ConstSpan*)plan.level_spans + it;
  __shared__ gaffa_hip_ffa_level_row sleaf[kLevelRowsMax];
  const uint32_t levels = item.a >> 16;
  // the host builds items within the level counts and the leaf tables
  if (item.c == 0 || item.c > kLevelRowsMax || levels < 1 || levels > 3) return;
  const gaffa_hip_ffa_task& task = plan.tasks[item.task];
  const uint64_t sidx = blockIdx.y;
  const uint32_t b = task.bins;
  const uint32_t depth = item.a & 0xFFFFu;
  const uint32_t nleaf = 1u << levels;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr uint32_t kWaves = kLevelThreads / 64;
  const uint64_t region = a.nseries * plan.work_stride;
  const float* src = a.work + static_cast<uint64_t>((task.regions >> (depth + levels)) & 1u) * region +
                     sidx * plan.work_stride + task.work_offset;
  float* dst = a.work + static_cast<uint64_t>((task.regions >> depth) & 1u) * region +
               sidx * plan.work_stride + task.work_offset;
  // the item's rows' leaves, to LDS beside the runs (read there per row)
  {
    const uint32_t* g = reinterpret_cast<const uint32_t*>(plan.level_leaf_rows + span->row_offset);
    const uint32_t total = item.c * (sizeof(gaffa_hip_ffa_level_row) / 4);
    uint32_t* l = reinterpret_cast<uint32_t*>(sleaf);
    for (uint32_t c0 = wave * 64u; c0 < total; c0 += kLevelThreads)
      if (c0 + lane < total) __builtin_amdgcn_global_load_lds(g + c0 + lane, l + c0, 4, 0, 0);
  }
  uint32_t run_lo[8], run_off[8];  // a leaf's first row, and its run's first staged row
  uint32_t staged = 0;
#pragma unroll
  for (uint32_t j = 0; j < 8; ++j) {
    run_lo[j] = 0;
    run_off[j] = 0;
    if (j < nleaf) {
      run_lo[j] = span->first[j];
      run_off[j] = staged;
      staged += span->count[j];
    }
  }
  const bool in_lds = staged == span->staged && level_stage_bytes(staged, b) <= plan.level_lds_bytes;
  if (in_lds) {
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
      if (j >= nleaf) break;
      const float* g = src + static_cast<uint64_t>(run_lo[j]) * b;
      float* l = stage + run_off[j] * b;
      const uint32_t total = span->count[j] * b;
      for (uint32_t c0 = wave * 64u; c0 < total; c0 += kLevelThreads)
        if (c0 + lane < total) __builtin_amdgcn_global_load_lds(g + c0 + lane, l + c0, 4, 0, 0);
    }
  }
  __builtin_amdgcn_s_waitcnt(0);  // the LDS writes of global_load_lds count in vmcnt
  __syncthreads();

  // out[p] for the row's phases, four 64-phase chunks per block.  From LDS,
  // each operand picks one of two addresses per chunk (before or after its
  // wrap point b - R_j) with one compare and one select, plus the chunk's
  // immediate offset; lanes past the row read inside the next staged row or
  // the guard.  From global memory, phases are clamped to the row and
  // wrapped one by one (the work region has no guard).
  auto merge_rows = [&](auto nlv, auto from_lds) {
    constexpr uint32_t NL = 1u << decltype(nlv)::value;
    constexpr bool kLds = decltype(from_lds)::value;
    using LdsFloat = __attribute__((address_space(3))) const float;
    for (uint32_t i = wave; i < item.c; i += kWaves) {
      LdsFloat* pa[NL];
      LdsFloat* pb[NL];
      const float* rp[NL];
      uint32_t roll[NL];
#pragma unroll
      for (uint32_t j = 0; j < NL; ++j) {
        roll[j] = __builtin_amdgcn_readfirstlane(sleaf[i].roll[j]);
        const uint32_t row = __builtin_amdgcn_readfirstlane(sleaf[i].row[j]);
        if constexpr (kLds) {
          pa[j] = (LdsFloat*)(stage + (run_off[j] + row - run_lo[j]) * b + roll[j] + lane);
          pb[j] = pa[j] - b;
        } else {
          rp[j] = src + static_cast<uint64_t>(row) * b;
        }
      }
      float* out = dst + static_cast<uint64_t>(item.b + i) * b;
      for (uint32_t p0 = 0; p0 < b; p0 += 256) {
        float v[NL][4];
#pragma unroll
        for (uint32_t j = 0; j < NL; ++j) {
          if constexpr (kLds) {
            const uint32_t wrap = b - roll[j];
            const uint32_t before = wrap - min(wrap, p0);  // phases of the block before the wrap
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
              LdsFloat* ptr = lane + 64u * k < before ? pa[j] : pb[j];
              v[j][k] = ptr[p0 + 64u * k];
            }
          } else {
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
              const uint32_t p = min(p0 + 64u * k + lane, b - 1);
              uint32_t x = p + roll[j];
              x = min(x, x - b);  // (p + R_j) mod b for p, R_j < b
              v[j][k] = rp[j][x];
            }
          }
        }
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
          float sum;
          if constexpr (NL == 2) {
            sum = v[0][k] + v[1][k];
          } else if constexpr (NL == 4) {
            sum = (v[0][k] + v[1][k]) + (v[2][k] + v[3][k]);
          } else {
            sum = ((v[0][k] + v[1][k]) + (v[2][k] + v[3][k])) +
                  ((v[4][k] + v[5][k]) + (v[6][k] + v[7][k]));
          }
          const uint32_t p = p0 + 64u * k + lane;
          if (p < b) __builtin_nontemporal_store(sum, &out[p]);  // streaming, as the subtree's
        }
      }
    }
  };
  auto merge = [&](auto from_lds) {
    if (levels == 1) merge_rows(std::integral_constant<uint32_t, 1>{}, from_lds);
    else if (levels == 2) merge_rows(std::integral_constant<uint32_t, 2>{}, from_lds);
    else merge_rows(std::integral_constant<uint32_t, 3>{}, from_lds);
  };
  if (in_lds) merge(std::true_type{});
  else merge(std::false_type{});
}

// ------------------------------------------------------------ terminal ----
__device__ __host__ __forceinline__ uint32_t terminal_stride(uint32_t bins, uint32_t maxw) {
  return (bins + maxw + 8) | 1u;  // odd: lane-per-row reads hit distinct banks
}

// max over phases p < b of P[p + w] - P[p] for one row x (circularly
// extended by at least w + 7 entries) with P the serial FP32 prefix;
// `total` = P[b].  The phases go in blocks of kScanBlock, and the next
// block's LDS reads are issued before the current block's adds: two register
// sets, the block loop unrolled by two, so that no set is copied into the
// other.  Blocks of 4 keep both blocks' 2 x 4 ds_read2 within the LDS
// counter's 15 outstanding operations, which blocks of 8 exceed.
constexpr uint32_t kScanBlock = 4;

__device__ __forceinline__ float boxcar_max(const float* x, uint32_t b, uint32_t w,
                                            float& total) {
  float A = 0.0f;  // P[p]
  float B = 0.0f;  // P[p + w]
  for (uint32_t k = 0; k < w; ++k) B += x[k];
  float best = -INFINITY;
  struct Block {
    float a[kScanBlock], b[kScanBlock];
  };
  auto load = [&](Block& v, uint32_t p) {
#pragma unroll
    for (uint32_t u = 0; u < kScanBlock; ++u) {
      v.a[u] = x[p + u];
      v.b[u] = x[p + w + u];
    }
  };
  auto step = [&](const Block& v) {
#pragma unroll
    for (uint32_t u = 0; u < kScanBlock; ++u) {
      best = fmaxf(best, B - A);
      A += v.a[u];
      B += v.b[u];
    }
  };
  const uint32_t nblocks = b / kScanBlock;
  if (nblocks > 0) {
    Block X, Y;
    load(X, 0);
    uint32_t k = 0;
    for (; k + 2 < nblocks; k += 2) {  // X holds block k
      load(Y, kScanBlock * (k + 1));
      step(X);
      load(X, kScanBlock * (k + 2));
      step(Y);
    }
    if (k + 1 < nblocks) {
      load(Y, kScanBlock * (k + 1));
      step(X);
      step(Y);
    } else {
      step(X);
    }
  }
  for (uint32_t p = kScanBlock * nblocks; p < b; ++p) {
    best = fmaxf(best, B - A);
    A += x[p];
    B += x[p + w];
  }
  total = A;
  return best;
}

template <bool kDetect>
__global__ void __launch_bounds__(kTerminalThreads)
ffa_terminal_kernel(KArgs a, float* __restrict__ rows_out, uint32_t only_task) {
  extern __shared__ float lds[];
  __shared__ Desc2 sdesc[kTerminalRowsMax];  // one-level merges use a, b, rh
  const gaffa_hip_ffa_plan_desc& plan = a.plan;
  const gaffa_hip_ffa_item item = plan.terminal_items[blockIdx.x];
  if (!kDetect && item.task != only_task) return;
  // by reference: the per-width tables are indexed at run time, and a
  // private copy of the 240-byte descriptor would live in scratch memory
  const gaffa_hip_ffa_task& task = plan.tasks[item.task];
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
  const uint32_t sb = terminal_stride(b, maxw);
  // the host sizes the LDS rows and the items together; an item beyond them
  // is refused rather than written past the tile
  if (nrows == 0 || nrows > kTerminalRowsMax ||
      static_cast<uint64_t>(nrows) * sb * 4u > plan.terminal_lds_bytes)
    return;

  // depth of the rows read: 1, or 2 when this kernel merges two levels
  const uint32_t tl = task.rows >= 2 && task.terminal_levels == 2 ? 2u : 1u;
  const float* src;
  if (task.rows == 1) {
    src = task_input(a, group, sidx);
  } else {
    src = a.work + static_cast<uint64_t>((task.regions >> tl) & 1u) * a.nseries * plan.work_stride +
          sidx * plan.work_stride +
          task.work_offset;
    if (threadIdx.x < nrows) {
      const uint32_t s = s0 + threadIdx.x;
      const Desc top = merge_desc(Node{0, task.rows}, s, b);
      if (tl == 1) {
        sdesc[threadIdx.x] = Desc2{top.h, top.t, top.roll, 0, 0, 0, 0};
      } else {
        const uint32_t hn = task.rows / 2;
        const Desc dh = merge_desc(Node{0, hn}, top.h, b);
        const Desc dt = merge_desc(Node{hn, task.rows - hn}, top.t, b);
        sdesc[threadIdx.x] = Desc2{dh.h, dh.t, dh.roll, dt.h, dt.t, dt.roll, top.roll};
      }
    }
    __syncthreads();
  }
  // A detection row x is stored with its circular extension x[q] = x[q mod
  // b] for q in [b, sb): x[p + b] together with x[p], and, only where
  // sb > 2b (widths near the bin count), q >= 2b in a pass of its own.
  auto put = [&](uint32_t i, uint32_t p, float v) {
    float* row = lds + static_cast<size_t>(i) * sb;
    row[p] = v;
    if constexpr (kDetect) {
      if (p + b < sb) row[p + b] = v;
    }
  };
  if (task.rows == 1) {
    for (uint32_t p = threadIdx.x; p < b; p += kTerminalThreads) put(0, p, src[p]);
  } else if (tl == 1) {
    constexpr uint32_t kBatch = 4;
    for (uint32_t i0 = wave; i0 < nrows; i0 += kBatch * kWaves) {
      merge_batch<kBatch>(
          src, i0, kWaves, nrows, b, lane,
          [&](uint32_t i, uint32_t& h, uint32_t& t, uint32_t& roll) {
            const Desc2 ds = sdesc[i];
            h = ds.a * b;
            t = ds.b * b;
            roll = ds.rh;
          },
          [&](uint32_t i, uint32_t p, float v) { put(i, p, v); });
    }
  } else {
    // two levels: out[p] = (A[p] + B[p+rh]) + (C[q] + D[q+rt]), q = p + ro
    // (indices mod b), the FP32 sums of the two one-level merges in order;
    // four rows per wave at a time, four phases per lane: 64 loads in flight
    constexpr uint32_t kBatch = 4;
    for (uint32_t i0 = wave; i0 < nrows; i0 += kBatch * kWaves) {
      Desc2 e[kBatch];
#pragma unroll
      for (uint32_t k = 0; k < kBatch; ++k) e[k] = sdesc[min(i0 + k * kWaves, nrows - 1)];
      for (uint32_t p0 = lane; p0 < b; p0 += 256) {
        float v[kBatch][4];
#pragma unroll
        for (uint32_t k = 0; k < kBatch; ++k) {
          // 32-bit element offsets from the uniform base (see ld32)
          const uint32_t A = e[k].a * b, B = e[k].b * b, C = e[k].c * b, D = e[k].d * b;
#pragma unroll
          for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t p = min(p0 + 64u * j, b - 1);
            uint32_t pb = p + e[k].rh;
            pb = min(pb, pb - b);
            uint32_t q = p + e[k].ro;
            q = min(q, q - b);
            uint32_t qd = q + e[k].rt;
            qd = min(qd, qd - b);
            v[k][j] = (ld32(src, A + p) + ld32(src, B + pb)) + (ld32(src, C + q) + ld32(src, D + qd));
          }
        }
#pragma unroll
        for (uint32_t k = 0; k < kBatch; ++k) {
          const uint32_t i = i0 + k * kWaves;
#pragma unroll
          for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t p = p0 + 64u * j;
            if (i < nrows && p < b) put(i, p, v[k][j]);
          }
        }
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
    if (sb > 2 * b) {  // uniform: the rest of the extension, from [0, b)
      for (uint32_t i = wave; i < nrows; i += kWaves) {
        float* row = lds + static_cast<size_t>(i) * sb;
        for (uint32_t q = 2 * b + lane; q < sb; q += 64) row[q] = row[q % b];
      }
      __syncthreads();
    }
    // One lane per (row, width) of this item: rows fastest.
    const uint32_t npairs = nrows * task.nwidths;
    for (uint32_t e = threadIdx.x; e < npairs; e += kTerminalThreads) {
      const uint32_t i = e % nrows;
      const uint32_t widx = e / nrows;
      const uint32_t w = task.width[widx];
      const float* x = lds + static_cast<size_t>(i) * sb;
      float total;
      const float best = boxcar_max(x, b, w, total);
      const float hb = task.height[widx] + task.baseline[widx];
      const float snr = (hb * best - task.baseline[widx] * total) * task.inv_stdnoise;
      if (a.periodogram != nullptr)
        a.periodogram[(sidx + task.series) * a.periodogram_stride +
                      static_cast<uint64_t>(task.periodogram_row + s0 + i) * a.periodogram_widths +
                      widx] = snr;
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
          pk.series = a.series_index_base + static_cast<uint32_t>(sidx) + task.series;
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
  check(a->periodogram == nullptr ||
            (a->periodogram_widths > 0 && a->periodogram_widths <= GAFFA_HIP_FFA_MAX_WIDTHS),
        "FFA periodogram width count must be in 1..16");
  check(p.max_bins > 1 && p.max_bins < 65536, "FFA bins must be in (1, 65536)");
  check(p.terminal_lds_bytes <= kMaxLdsBytes, "FFA detection tile exceeds LDS");
  check(p.level_lds_bytes <= kMaxLevelLdsBytes, "FFA level staging exceeds LDS");
  check(p.level_item_offsets == nullptr || p.level_item_offsets[p.max_depth] == 0 ||
            (p.level_spans != nullptr && p.level_leaf_rows != nullptr),
        "FFA level items need their spans and leaf rows");
  check(p.subtree_lds_bytes <= kMaxLdsBytes, "FFA subtree exceeds LDS");
  check(p.nsubtree_items == 0 ||
            (p.subtree_half_floats > 0 &&
             p.subtree_half_floats <= kSubtreePrefetch * kSubtreeThreads &&
             (p.subtree_merges != nullptr || p.max_subtree_desc == 0) &&
             8ull * p.subtree_half_floats + 4ull * kSubtreeGuard +
                     4ull * subtree_zero_floats(p.max_bins) <= p.subtree_lds_bytes),
        "FFA subtree geometry inconsistent");
}

}  // namespace
}  // namespace gaffa_hip

using namespace gaffa_hip;

namespace {

// Raises a kernel's dynamic-LDS limit to at least `bytes`, never lowering
// it: launches captured into a graph replay under the largest limit any of
// them needed.  Per device and kernel (slot): subtree, detection, transform
// rows, level.
void ensure_dynamic_lds(const void* kernel, int slot, uint32_t bytes) {
  static uint32_t set[64][4] = {};
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  check(dev >= 0 && dev < 64, "device ordinal out of range");
  if (bytes <= set[dev][slot]) return;
  hip_check(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                static_cast<int>(bytes)),
            "hipFuncSetAttribute");
  set[dev][slot] = bytes;
}

int compute_units() {
  // per device, queried once
  static int cus[64] = {};
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  check(dev >= 0 && dev < 64, "device ordinal out of range");
  if (cus[dev] == 0) {
    int n = 0;
    hip_check(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev),
              "hipDeviceGetAttribute");
    cus[dev] = n > 0 ? n : 1;
  }
  return cus[dev];
}

void launch_prepare(const gaffa_hip_ffa_search_args* a, hipStream_t st) {
  const gaffa_hip_ffa_plan_desc& p = *a->plan;
  if (p.ngroups == 0 || p.max_prepared_nsamples == 0) return;
  check(a->prepared != nullptr, "FFA prepared buffer must not be null");
  check(p.prep_blocks > 0 && p.prep_blocks < (1u << 31), "FFA prepare geometry missing");
  if (p.prep_fused_halo > 0) {
    check(p.prep_fused_halo <= kPrepMaxHalo && p.ngroups <= kPrepMaxGroups &&
              p.prep_input_nsamples > 0,
          "FFA fused prepare geometry out of range");
    const dim3 grid(static_cast<uint32_t>(ceil_div(p.prep_input_nsamples, kPrepSpan)), 1,
                    static_cast<uint32_t>(a->nseries));
    const uint32_t lds = (kPrepSpan + p.prep_fused_halo) * 4u;
    hipLaunchKernelGGL(ffa_prepare_fused_kernel, grid, dim3(kPrepThreads), lds, st, a->series,
                       a->series_stride, p.groups, p.ngroups, p.prep_fused_halo, a->prepared,
                       p.prepared_stride);
    launch_check("ffa_prepare_fused_kernel");
    return;
  }
  // groups needing fewer blocks than the widest return at once
  const dim3 grid(p.prep_blocks, p.ngroups, static_cast<uint32_t>(a->nseries));
  hipLaunchKernelGGL(ffa_prepare_kernel, grid, dim3(kPrepThreads), 0, st, a->series,
                     a->series_stride, p.groups, a->prepared, p.prepared_stride);
  launch_check("ffa_prepare_kernel");
}

void launch_subtrees(const gaffa_hip_ffa_search_args* a, hipStream_t st) {
  const gaffa_hip_ffa_plan_desc& p = *a->plan;
  const uint32_t ns = static_cast<uint32_t>(a->nseries);
  if (p.nsubtree_items > 0) {
    ensure_dynamic_lds(reinterpret_cast<const void*>(ffa_subtree_kernel), 0, p.subtree_lds_bytes);
    const uint64_t nunits = static_cast<uint64_t>(p.nsubtree_items) * ns;
    check_overflow(nunits < (1ull << 32), "FFA subtree work too large");
    // one resident workgroup per CU (the LDS footprint allows no more)
    const uint32_t grid = static_cast<uint32_t>(
        std::min<uint64_t>(nunits, static_cast<uint64_t>(compute_units())));
    hipLaunchKernelGGL(ffa_subtree_kernel, dim3(grid), dim3(kSubtreeThreads),
                       p.subtree_lds_bytes, st, kargs(a), static_cast<uint32_t>(nunits));
    launch_check("ffa_subtree_kernel");
  }
}

void launch_levels(const gaffa_hip_ffa_search_args* a, hipStream_t st) {
  const gaffa_hip_ffa_plan_desc& p = *a->plan;
  const uint32_t ns = static_cast<uint32_t>(a->nseries);
  for (int32_t d = static_cast<int32_t>(p.max_depth) - 1; d >= 1; --d) {
    const uint32_t lo = p.level_item_offsets[d];
    const uint32_t hi = p.level_item_offsets[d + 1];
    if (hi <= lo) continue;
    ensure_dynamic_lds(reinterpret_cast<const void*>(ffa_level_kernel), 3, p.level_lds_bytes);
    hipLaunchKernelGGL(ffa_level_kernel, dim3(hi - lo, ns), dim3(kLevelThreads), p.level_lds_bytes,
                       st, kargs(a), lo);
    launch_check("ffa_level_kernel");
  }
}

void launch_transform(const gaffa_hip_ffa_search_args* a, hipStream_t st) {
  launch_subtrees(a, st);
  launch_levels(a, st);
}

void validate_search(const gaffa_hip_ffa_search_args* a) {
  validate(a);
  check(a->peaks != nullptr && a->peak_count != nullptr, "FFA peak buffers must not be null");
  check(std::isfinite(a->snr_threshold), "FFA search S/N threshold must be finite");
}

template <bool kDetect>
void launch_terminal(const gaffa_hip_ffa_search_args* a, float* out, uint32_t task,
                     hipStream_t st) {
  const gaffa_hip_ffa_plan_desc& p = *a->plan;
  ensure_dynamic_lds(reinterpret_cast<const void*>(ffa_terminal_kernel<kDetect>), kDetect ? 1 : 2,
                     p.terminal_lds_bytes);
  hipLaunchKernelGGL(ffa_terminal_kernel<kDetect>,
                     dim3(p.nterminal_items, static_cast<uint32_t>(a->nseries)),
                     dim3(kTerminalThreads), p.terminal_lds_bytes, st, kargs(a), out, task);
  launch_check("ffa_terminal_kernel");
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
    validate_search(a);
    const hipStream_t st = as_stream(stream);
    launch_prepare(a, st);
    launch_transform(a, st);
    launch_terminal<true>(a, nullptr, 0, st);
  });
}

int gaffa_hip_ffa_subtree_batch(const gaffa_hip_ffa_search_args* a, gaffa_hip_stream stream) {
  return guarded([&] {
    validate(a);
    launch_subtrees(a, as_stream(stream));
  });
}

int gaffa_hip_ffa_levels_batch(const gaffa_hip_ffa_search_args* a, gaffa_hip_stream stream) {
  return guarded([&] {
    validate(a);
    launch_levels(a, as_stream(stream));
  });
}

int gaffa_hip_ffa_detect_batch(const gaffa_hip_ffa_search_args* a, gaffa_hip_stream stream) {
  return guarded([&] {
    validate_search(a);
    launch_terminal<true>(a, nullptr, 0, as_stream(stream));
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
    launch_terminal<false>(a, out, task, st);
  });
}

int gaffa_hip_ffa_reserve_launch(const gaffa_hip_ffa_plan_desc* p) {
  return guarded([&] {
    check(p != nullptr, "FFA plan must not be null");
    check(p->subtree_lds_bytes <= kMaxLdsBytes && p->terminal_lds_bytes <= kMaxLdsBytes &&
              p->level_lds_bytes <= kMaxLevelLdsBytes,
          "FFA plan exceeds LDS");
    if (p->nsubtree_items > 0)
      ensure_dynamic_lds(reinterpret_cast<const void*>(ffa_subtree_kernel), 0,
                         p->subtree_lds_bytes);
    ensure_dynamic_lds(reinterpret_cast<const void*>(ffa_terminal_kernel<true>), 1,
                       p->terminal_lds_bytes);
    ensure_dynamic_lds(reinterpret_cast<const void*>(ffa_level_kernel), 3, p->level_lds_bytes);
    (void)compute_units();
  });
}

uint32_t gaffa_hip_ffa_prepare_outputs_per_block(double factor) {
  return prep_outputs_per_block(factor);
}

uint32_t gaffa_hip_ffa_terminal_rows(void) { return kTerminalRows; }
uint32_t gaffa_hip_ffa_level_rows(void) { return kLevelRows; }
uint32_t gaffa_hip_ffa_terminal_rows_max(void) { return kTerminalRowsMax; }
uint32_t gaffa_hip_ffa_level_rows_max(void) { return kLevelRowsMax; }

uint64_t gaffa_hip_ffa_level_item(uint32_t task_rows, uint32_t bins, uint32_t depth,
                                  uint32_t levels, uint32_t first, uint32_t count,
                                  gaffa_hip_ffa_level_span* span, gaffa_hip_ffa_level_row* rows) {
  if (levels < 1 || levels > 3 || bins < 2 || count == 0 || count > kLevelRowsMax ||
      task_rows < 2 || first >= task_rows || count > task_rows - first)
    return ~uint64_t{0};
  const uint32_t nleaf = 1u << levels;
  uint32_t lo[8] = {}, hi[8] = {};
  for (uint32_t i = 0; i < count; ++i) {
    const gaffa_hip_ffa_level_row e = level_leaves(task_rows, depth, levels, first + i, bins);
    if (rows != nullptr) rows[i] = e;
    for (uint32_t j = 0; j < nleaf; ++j) {
      lo[j] = i == 0 ? e.row[j] : std::min(lo[j], e.row[j]);
      hi[j] = i == 0 ? e.row[j] : std::max(hi[j], e.row[j]);
    }
  }
  uint32_t staged = 0;
  for (uint32_t j = 0; j < nleaf; ++j) staged += hi[j] - lo[j] + 1;
  if (span != nullptr) {
    span->staged = staged;
    for (uint32_t j = 0; j < 8; ++j) {
      span->first[j] = j < nleaf ? lo[j] : 0;
      span->count[j] = j < nleaf ? hi[j] - lo[j] + 1 : 0;
    }
    span->reserved[0] = span->reserved[1] = 0;
  }
  return level_stage_bytes(staged, bins);
}

uint64_t gaffa_hip_ffa_level_lds_bytes(uint32_t task_rows, uint32_t bins, uint32_t depth,
                                       uint32_t levels, uint32_t first, uint32_t count) {
  return gaffa_hip_ffa_level_item(task_rows, bins, depth, levels, first, count, nullptr, nullptr);
}

uint32_t gaffa_hip_ffa_terminal_lds_bytes(uint32_t bins, uint32_t max_width) {
  return kTerminalRows * terminal_stride(bins, max_width) * 4u;
}

uint32_t gaffa_hip_ffa_terminal_lds_bytes_rows(uint32_t bins, uint32_t max_width, uint32_t rows) {
  if (rows == 0 || rows > kTerminalRowsMax) return 0xFFFFFFFFu;
  return rows * terminal_stride(bins, max_width) * 4u;
}

uint32_t gaffa_hip_ffa_subtree_lds_bytes(uint32_t rows, uint32_t bins) {
  // two row buffers, the zero rows in front of them and the slack after
  if (rows == 0 || rows > kSubtreeMaxRows ||
      static_cast<uint64_t>(rows) * bins > kSubtreePrefetch * kSubtreeThreads)
    return 0xFFFFFFFFu;
  return 8u * rows * bins + 4u * kSubtreeGuard + 4u * subtree_zero_floats(bins);
}

uint32_t gaffa_hip_ffa_subtree_merges(uint32_t rows, uint32_t bins, gaffa_hip_ffa_merge* out) {
  if (rows == 0 || rows > kSubtreeMaxRows || bins < 2 || bins > 65535) return 0;
  const uint32_t levels = tree_levels(rows);
  const uint32_t npasses = subtree_tree_passes(levels);
  const uint32_t per_wave = kSubtreeRowsPerPass * subtree_passes(rows);
  const uint32_t pass_size = kSubtreeWaves * per_wave;
  if (out == nullptr) return npasses * pass_size;
  constexpr uint32_t kCopy = GAFFA_HIP_FFA_COPY_ROW;
  // row r of the merge of the node containing r at `depth`, read from the
  // rows one level below: head row, tail row and roll, or a copied leaf
  auto one = [&](uint32_t depth, uint32_t r, uint32_t& head, uint32_t& tail, uint32_t& roll) {
    const Node nd = find_node(0, rows, depth, r);
    if (nd.size < 2) {
      head = r;
      tail = kCopy;
      roll = 0;
    } else {
      const Desc ds = merge_desc(nd, r, bins);
      head = ds.h;
      tail = ds.t;
      roll = ds.roll;
    }
  };
  for (uint32_t k = 0; k < npasses; ++k) {
    const uint32_t depth = 2 * k;                      // output depth of the pass
    const bool two = depth + 1 < levels;               // else the bottom single level
    for (uint32_t w = 0; w < kSubtreeWaves; ++w) {
      for (uint32_t j = 0; j < per_wave; ++j) {
        // row w + 16 j, or padding past the wave's last row
        const uint32_t r = w + kSubtreeWaves * j;
        gaffa_hip_ffa_merge& m = out[k * pass_size + w * per_wave + j];
        m = gaffa_hip_ffa_merge{0, kCopy, bins, GAFFA_HIP_FFA_SKIP_ROW, kCopy, bins, kCopy, bins};
        if (r >= rows) continue;
        m.out = r * bins;
        uint32_t h, t, roll;
        one(depth, r, h, t, roll);
        if (!two) {
          m.head = h * bins;
          if (t != kCopy) {
            m.tail = t * bins + roll;
            m.wrap = bins - roll;
          }
          continue;
        }
        // head child row h: A + B; tail child row t (read at `roll`): C + D
        uint32_t a, b, rb;
        one(depth + 1, h, a, b, rb);
        m.head = a * bins;
        if (b != kCopy) {
          m.tail = b * bins + rb;
          m.wrap = bins - rb;
        }
        if (t == kCopy) continue;  // a leaf above depth: (A + -0) + (-0 + -0)
        uint32_t c, d, rd;
        one(depth + 1, t, c, d, rd);
        m.head2 = c * bins + roll;
        m.wrap2 = bins - roll;
        if (d != kCopy) {
          const uint32_t rcd = (roll + rd) % bins;
          m.tail2 = d * bins + rcd;
          m.wrap3 = bins - rcd;
        }
      }
    }
  }
  return npasses * pass_size;
}

}  // extern "C"
