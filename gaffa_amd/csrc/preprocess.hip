// Batched in-place preprocessing of a DM x time plane on gfx950: uint32 ->
// float conversion, riptide's fast running-median detrend and normalisation
// (preprocessing.cpp:53-145, 212-231; time_series.cpp:50-104).
//
// Kernels (one launch each, all series at once, grid.y = series):
//   scrunch   - low[k] = float(sum_j double(x[k*f + j]) / f), serial FP64 sum
//               per output so the rounding equals the CPU loop
//   median    - exact centred median of an edge-padded window: a workgroup
//               bitonic-sorts (value, padded position) pairs of its tile's
//               union window in LDS, then each output counts members of its
//               own window in sorted order
//   detrend   - FP64 linear interpolation of the low-resolution median back to
//               full resolution (same operation order as the CPU), subtract,
//               write float
//   stats     - one lane per series: the FP64 sum and sum of squares in
//               sample order, exactly the CPU loop's roundings, then mean and
//               1/stddev (time_series.cpp:50-77)
//   normalise - (x - mean) * (1/stddev) in FP64, rounded to float
// Integer-valued inputs (dedispersed sums) make the scrunch sums exact, and
// every other FP operation runs in the CPU's order, so the output is the
// CPU's bit for bit.
#include "common.hpp"

#include <mutex>

#include <cmath>
#include <type_traits>

namespace gaffa_hip {
namespace {

constexpr int kThreads = 256;
constexpr int kDetrendItems = 8;  // samples per thread in detrend/normalise
constexpr int kBlockSamples = kThreads * kDetrendItems;

template <typename T>
__device__ inline float load_as_float(const T* p, uint64_t i) {
  return static_cast<float>(p[i]);
}

template <typename T>
__global__ void __launch_bounds__(kThreads)
scrunch_kernel(const T* __restrict__ data, uint64_t stride, uint64_t factor,
               uint64_t m, float* __restrict__ low) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  const uint64_t series = blockIdx.y;
  if (k >= m) return;
  const T* x = data + series * stride + k * factor;
  double sum = 0.0;
  for (uint64_t j = 0; j < factor; ++j) sum += static_cast<double>(load_as_float(x, j));
  low[series * m + k] = static_cast<float>(sum / static_cast<double>(factor));
}

// uint32 input: each sample is rounded to float (as gaffa converts a uint32
// DM series), which leaves an integer of at most 2^32; the sum of f of them
// is exact in a 64-bit integer, hence equal to the CPU's serial FP64 sum.  One
// wave per output, coalesced loads, shuffle reduction.
__global__ void __launch_bounds__(kThreads)
scrunch_u32_kernel(const uint32_t* __restrict__ data, uint64_t stride, uint64_t factor,
                   uint64_t m, float* __restrict__ low) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * (kThreads / 64) + (threadIdx.x >> 6);
  const uint64_t series = blockIdx.y;
  if (k >= m) return;  // wave-uniform
  const uint32_t* x = data + series * stride + k * factor;
  uint64_t sum = 0;
  for (uint64_t j = lane; j < factor; j += 64)
    sum += static_cast<uint64_t>(static_cast<float>(x[j]));
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
  if (lane == 0)
    low[series * m + k] = static_cast<float>(static_cast<double>(sum) / static_cast<double>(factor));
}

// Median of every centred window of `window` (odd) points over an
// edge-padded sequence of length m.  SORT >= kThreads + window - 1.
template <int SORT, typename T>
__global__ void __launch_bounds__(kThreads)
median_kernel(const T* __restrict__ src, uint64_t src_stride, uint64_t m,
              uint32_t window, float* __restrict__ med, uint64_t med_stride) {
  __shared__ float key[SORT];
  __shared__ int32_t pos[SORT];
  const uint32_t tid = threadIdx.x;
  const uint64_t series = blockIdx.y;
  const int64_t k0 = static_cast<int64_t>(blockIdx.x) * kThreads;
  const int32_t half = static_cast<int32_t>(window / 2);
  const int32_t span = kThreads + static_cast<int32_t>(window) - 1;
  const T* s = src + series * src_stride;
  for (int32_t e = tid; e < SORT; e += kThreads) {
    if (e < span) {
      const int64_t p = k0 - half + e;  // padded position
      const int64_t idx = p < 0 ? 0 : (p >= static_cast<int64_t>(m) ? static_cast<int64_t>(m) - 1 : p);
      key[e] = load_as_float(s, static_cast<uint64_t>(idx));
      pos[e] = e;
    } else {
      key[e] = INFINITY;
      pos[e] = 0x7fffffff;
    }
  }
  __syncthreads();
  // bitonic sort ascending by key
  for (int32_t size = 2; size <= SORT; size <<= 1) {
    for (int32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (int32_t e = tid; e < SORT / 2; e += kThreads) {
        const int32_t lo = 2 * e - (e & (stride - 1));
        const int32_t hi = lo + stride;
        const bool ascending = (lo & size) == 0;
        const float a = key[lo], b = key[hi];
        if ((a > b) == ascending) {
          key[lo] = b;
          key[hi] = a;
          const int32_t t = pos[lo];
          pos[lo] = pos[hi];
          pos[hi] = t;
        }
      }
      __syncthreads();
    }
  }
  const int64_t k = k0 + tid;
  if (k >= static_cast<int64_t>(m)) return;
  // window of output tid covers padded offsets [tid, tid + window)
  const uint32_t lo = tid;
  int32_t count = 0;
  float value = 0.0f;
  for (int32_t e = 0; e < span; ++e) {
    const uint32_t off = static_cast<uint32_t>(pos[e]) - lo;
    if (off < window) {
      if (count == half) {
        value = key[e];
        break;
      }
      ++count;
    }
  }
  med[series * med_stride + static_cast<uint64_t>(k)] = value;
}

// x - interpolated median, written in place as float.
template <typename T>
__global__ void __launch_bounds__(kThreads)
detrend_kernel(T* __restrict__ data, uint64_t stride, uint64_t n,
               int32_t detrend, uint64_t factor, const float* __restrict__ med,
               uint64_t m, int32_t* __restrict__ status) {
  const uint32_t tid = threadIdx.x;
  const uint64_t series = blockIdx.y;
  T* x = data + series * stride;
  float* out = reinterpret_cast<float*>(x);
  const float* md = med + series * (factor == 1 ? n : m);
  const double offset = 0.5 * static_cast<double>(factor - 1);
  const double fd = static_cast<double>(factor);
  bool bad = false;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlockSamples;
#pragma unroll
  for (int r = 0; r < kDetrendItems; ++r) {
    const uint64_t i = base + static_cast<uint64_t>(r) * kThreads + tid;
    if (i >= n) break;
    const float v = load_as_float(x, i);
    bad |= !isfinite(v);
    float d = v;
    if (detrend) {
      float baseline;
      if (factor == 1) {
        baseline = md[i];
      } else {
        const double p = (static_cast<double>(i) - offset) / fd;
        if (p <= 0.0) {
          baseline = md[0];
        } else {
          const uint64_t lo = static_cast<uint64_t>(floor(p));
          if (lo >= m - 1) {
            baseline = md[m - 1];
          } else {
            const double fr = p - static_cast<double>(lo);
            const double w0 = (1.0 - fr) * static_cast<double>(md[lo]);
            const double w1 = fr * static_cast<double>(md[lo + 1]);
            baseline = static_cast<float>(w0 + w1);
          }
        }
      }
      d = v - baseline;
    }
    out[i] = d;
  }
  if (bad) atomicExch(&status[series], 1);
}

// Sum and sum of squares of each series in FP64 in sample order, as
// time_series_stats_cpu does (time_series.cpp:55-60), so every rounding is
// the CPU's; then mean, 1/stddev and a status flag.
//
// A two-wave workgroup owns 64 series, a lane of each wave per series: wave 0
// runs the chain of sums, wave 1 the chain of squares (the two are
// independent, so each wave has one serial chain instead of two and the
// launch twice the waves).  The workgroup loads chunks of 64 series x 128
// samples with coalesced 16-byte loads (two 512-byte row segments per
// instruction) into registers, stores them to LDS, and while each lane adds
// its own row of the chunk in order, the next chunk's loads are already in
// flight.  Rows must be 16-byte aligned (the pipeline's plane is); other
// layouts take the per-lane path, which reads each row directly.
// a native 4-float vector: HIP's float4 struct is copied through memory,
// which would keep a register array of them in scratch
typedef float vf4 __attribute__((ext_vector_type(4)));
constexpr int kStatsLanes = 64;                // series per workgroup
constexpr int kStatsThreads = 2 * kStatsLanes; // one wave per chain
constexpr int kStatsChunk = 128;               // samples per series per chunk
constexpr int kStatsPitch = kStatsChunk + 4;   // LDS floats per row (bank spread)
constexpr int kStatsLoads = kStatsLanes * kStatsChunk / 4 / kStatsThreads;  // float4 per thread
constexpr uint32_t kStatsVecs = kStatsChunk / 4;  // float4 per row of a chunk
constexpr uint32_t kStatsLdsBytes = 2u * kStatsLanes * kStatsPitch * 4u;

struct SerialSums {
  double sum = 0.0, squares = 0.0;
  __device__ __forceinline__ void add(float v) {
    const double x = static_cast<double>(v);
    sum = sum + x;
    squares = squares + x * x;  // separately rounded: the code is built without contraction
  }
  __device__ __forceinline__ void add4(float4 v) {
    add(v.x);
    add(v.y);
    add(v.z);
    add(v.w);
  }
};

__device__ __forceinline__ void finish_stats(const SerialSums& acc, uint64_t n, uint32_t series,
                                             int32_t reject_constant, double2* stats,
                                             int32_t* flags, int32_t* status) {
  const double count = static_cast<double>(n);
  const double mean = acc.sum / count;
  double var = acc.squares / count - mean * mean;
  if (var < 0.0 && var > -1.0e-12) var = 0.0;
  int32_t flag = 0;
  if (var < 0.0 || !isfinite(var)) flag = 3;
  const double sd = sqrt(var);
  if (flag == 0 && (!(sd > 0.0) || !isfinite(sd))) flag = reject_constant ? 2 : 4;
  stats[series] = make_double2(mean, flag == 0 ? 1.0 / sd : 0.0);
  flags[series] = flag;
  if (flag == 2 || flag == 3) atomicCAS(&status[series], 0, flag);
}

__global__ void __launch_bounds__(kStatsThreads) __attribute__((amdgpu_waves_per_eu(1, 2)))
stats_kernel(const float* __restrict__ data, uint64_t stride, uint64_t n, uint32_t nseries,
             int32_t reject_constant, double2* __restrict__ stats,
             int32_t* __restrict__ flags, int32_t* __restrict__ status) {
  extern __shared__ __align__(16) float stage[];  // [2][kStatsLanes][kStatsPitch]
  __shared__ double squares_of[kStatsLanes];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const bool squares = tid >= 64u;  // wave-uniform: wave 1 runs the squares
  const uint32_t first = blockIdx.x * kStatsLanes;
  const uint32_t rows = min(static_cast<uint32_t>(kStatsLanes), nseries - first);
  const uint64_t nchunks = n / kStatsChunk;
  const float* base = data + static_cast<uint64_t>(first) * stride;
  // load k of a chunk: element e = k * 128 + tid is row e / kStatsVecs,
  // float4 e % kStatsVecs
  // (rows past the batch re-read the last row); every index is a constant
  // after unrolling, so the chunk stays in registers
  vf4 reg[kStatsLoads];
  double acc = 0.0;
  if (nchunks > 0) {
#pragma unroll
    for (int k = 0; k < kStatsLoads; ++k) {
      const uint32_t e = static_cast<uint32_t>(k) * kStatsThreads + tid;
      const uint32_t r = min(e / kStatsVecs, rows - 1);
      reg[k] = *reinterpret_cast<const vf4*>(base + static_cast<uint64_t>(r) * stride +
                                                4u * (e % kStatsVecs));
    }
  }
  for (uint64_t c = 0; c < nchunks; ++c) {
    const uint32_t buf = static_cast<uint32_t>(c & 1u);
#pragma unroll
    for (int k = 0; k < kStatsLoads; ++k) {
      const uint32_t e = static_cast<uint32_t>(k) * kStatsThreads + tid;
      *reinterpret_cast<vf4*>(stage + (buf * kStatsLanes + e / kStatsVecs) * kStatsPitch +
                                 4u * (e % kStatsVecs)) = reg[k];
    }
    __syncthreads();
    if (c + 1 < nchunks) {  // in flight while this chunk is summed
      const uint64_t col = (c + 1) * kStatsChunk;
#pragma unroll
      for (int k = 0; k < kStatsLoads; ++k) {
        const uint32_t e = static_cast<uint32_t>(k) * kStatsThreads + tid;
        const uint32_t r = min(e / kStatsVecs, rows - 1);
        reg[k] = *reinterpret_cast<const vf4*>(base + static_cast<uint64_t>(r) * stride + col +
                                                  4u * (e % kStatsVecs));
      }
    }
    const vf4* row = reinterpret_cast<const vf4*>(stage + (buf * kStatsLanes + lane) *
                                                  kStatsPitch);
    // eight samples at a time: the conversions (and squares) first,
    // independent of each other, then the in-order add chain, which is all
    // that remains on the critical path
    if (!squares) {
#pragma unroll 2
      for (int j = 0; j < kStatsChunk / 4; j += 2) {
        const vf4 v0 = row[j], v1 = row[j + 1];
        const float f[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        double d[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) d[u] = static_cast<double>(f[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = acc + d[u];
      }
    } else {
#pragma unroll 2
      for (int j = 0; j < kStatsChunk / 4; j += 2) {
        const vf4 v0 = row[j], v1 = row[j + 1];
        const float f[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
        double q[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const double d = static_cast<double>(f[u]);
          q[u] = d * d;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) acc = acc + q[u];
      }
    }
    // the other buffer is rewritten after the next barrier, once both waves
    // are done with it
  }
  // the tail (< 128 samples) of each row, then the squares join the sums
  if (lane < rows) {
    const float* x = base + static_cast<uint64_t>(lane) * stride;
    for (uint64_t i = nchunks * kStatsChunk; i < n; ++i) {
      const double d = static_cast<double>(x[i]);
      acc = squares ? acc + d * d : acc + d;
    }
    if (squares) squares_of[lane] = acc;
  }
  __syncthreads();
  if (squares || lane >= rows) return;
  SerialSums sums;
  sums.sum = acc;
  sums.squares = squares_of[lane];
  finish_stats(sums, n, first + lane, reject_constant, stats, flags, status);
}

// Rows that are not 16-byte aligned: each lane reads its row directly.
__global__ void __launch_bounds__(kStatsLanes)
stats_rowwise_kernel(const float* __restrict__ data, uint64_t stride, uint64_t n,
                     uint32_t nseries, int32_t reject_constant, double2* __restrict__ stats,
                     int32_t* __restrict__ flags, int32_t* __restrict__ status) {
  const uint32_t series = blockIdx.x * kStatsLanes + threadIdx.x;
  if (series >= nseries) return;
  const float* x = data + static_cast<uint64_t>(series) * stride;
  SerialSums acc;
  for (uint64_t i = 0; i < n; ++i) acc.add(x[i]);
  finish_stats(acc, n, series, reject_constant, stats, flags, status);
}

__global__ void __launch_bounds__(kThreads)
normalise_kernel(float* __restrict__ data, uint64_t stride, uint64_t n,
                 const double2* __restrict__ stats, const int32_t* __restrict__ flags) {
  const uint32_t tid = threadIdx.x;
  const uint64_t series = blockIdx.y;
  const int32_t flag = flags[series];
  if (flag == 2 || flag == 3) return;  // rejected: data left detrended, status set
  const double mean = stats[series].x, scale = stats[series].y;
  float* x = data + series * stride;
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kBlockSamples;
#pragma unroll
  for (int r = 0; r < kDetrendItems; ++r) {
    const uint64_t i = base + static_cast<uint64_t>(r) * kThreads + tid;
    if (i >= n) break;
    // flag 4: constant input accepted -> zeros (time_series.cpp:90-93)
    x[i] = flag == 4 ? 0.0f
                     : static_cast<float>((static_cast<double>(x[i]) - mean) * scale);
  }
}

uint64_t median_points(uint64_t n, uint64_t factor) { return factor == 1 ? n : n / factor; }

template <typename T>
void launch_median(const T* src, uint64_t src_stride, uint64_t m, uint32_t window,
                   float* med, uint64_t med_stride, uint64_t nseries, hipStream_t st) {
  const dim3 grid(static_cast<uint32_t>(ceil_div(m, kThreads)), static_cast<uint32_t>(nseries));
  const uint32_t span = kThreads + window - 1;
  if (span <= 512) {
    hipLaunchKernelGGL((median_kernel<512, T>), grid, dim3(kThreads), 0, st, src, src_stride, m,
                       window, med, med_stride);
  } else if (span <= 1024) {
    hipLaunchKernelGGL((median_kernel<1024, T>), grid, dim3(kThreads), 0, st, src, src_stride, m,
                       window, med, med_stride);
  } else if (span <= 2048) {
    hipLaunchKernelGGL((median_kernel<2048, T>), grid, dim3(kThreads), 0, st, src, src_stride, m,
                       window, med, med_stride);
  } else {
    hipLaunchKernelGGL((median_kernel<4096, T>), grid, dim3(kThreads), 0, st, src, src_stride, m,
                       window, med, med_stride);
  }
  launch_check("median_kernel");
}

struct Layout {
  uint64_t m, low_floats, med_floats, nblocks, stats_offset,
      flags_offset, total;
};

Layout layout(uint64_t nseries, uint64_t n, uint64_t factor) {
  Layout l{};
  l.m = median_points(n, factor);
  l.low_floats = factor == 1 ? 0 : nseries * l.m;
  l.med_floats = nseries * l.m;
  l.nblocks = ceil_div(n, kBlockSamples);
  l.stats_offset = ((l.low_floats + l.med_floats) * 4 + 255) / 256 * 256;
  l.flags_offset = l.stats_offset + (nseries * sizeof(double2) + 255) / 256 * 256;
  l.total = l.flags_offset + nseries * sizeof(int32_t);
  return l;
}

template <typename T>
void run(const gaffa_hip_preprocess_args* a, void* workspace, hipStream_t st) {
  const Layout l = layout(a->nseries, a->nsamples, a->detrend ? a->scrunch_factor : 1);
  float* low = static_cast<float*>(workspace);
  float* med = low + l.low_floats;
  char* ws = static_cast<char*>(workspace);
  double2* stats = reinterpret_cast<double2*>(ws + l.stats_offset);
  int32_t* flags = reinterpret_cast<int32_t*>(ws + l.flags_offset);
  T* data = static_cast<T*>(a->data);
  const uint64_t f = a->scrunch_factor;
  if (a->detrend) {
    if (f > 1) {
      if constexpr (std::is_same_v<T, uint32_t>) {
        const dim3 g(static_cast<uint32_t>(ceil_div(l.m, kThreads / 64)),
                     static_cast<uint32_t>(a->nseries));
        hipLaunchKernelGGL(scrunch_u32_kernel, g, dim3(kThreads), 0, st, data, a->stride, f, l.m,
                           low);
      } else {
        const dim3 g(static_cast<uint32_t>(ceil_div(l.m, kThreads)),
                     static_cast<uint32_t>(a->nseries));
        hipLaunchKernelGGL((scrunch_kernel<T>), g, dim3(kThreads), 0, st, data, a->stride, f,
                           l.m, low);
      }
      launch_check("scrunch_kernel");
      launch_median<float>(low, l.m, l.m, static_cast<uint32_t>(a->median_window), med, l.m,
                           a->nseries, st);
    } else {
      launch_median<T>(data, a->stride, a->nsamples, static_cast<uint32_t>(a->median_window), med,
                       a->nsamples, a->nseries, st);
    }
  }
  const dim3 g(static_cast<uint32_t>(l.nblocks), static_cast<uint32_t>(a->nseries));
  hipLaunchKernelGGL((detrend_kernel<T>), g, dim3(kThreads), 0, st, data, a->stride, a->nsamples,
                     a->detrend, a->detrend ? f : 1, med, l.m, a->status);
  launch_check("detrend_kernel");
  if (a->normalise) {
    const dim3 sg(static_cast<uint32_t>(ceil_div(a->nseries, kStatsLanes)));
    const float* fdata = reinterpret_cast<const float*>(data);
    if ((reinterpret_cast<uintptr_t>(fdata) & 15u) == 0 && a->stride % 4 == 0) {
      // two chunk buffers exceed the default 64 KiB of dynamic LDS
      hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(stats_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    static_cast<int>(kStatsLdsBytes)),
                "hipFuncSetAttribute");
      hipLaunchKernelGGL(stats_kernel, sg, dim3(kStatsThreads), kStatsLdsBytes, st, fdata,
                         a->stride, a->nsamples, static_cast<uint32_t>(a->nseries),
                         a->reject_constant, stats, flags, a->status);
      launch_check("stats_kernel");
    } else {
      hipLaunchKernelGGL(stats_rowwise_kernel, sg, dim3(kStatsLanes), 0, st, fdata, a->stride,
                         a->nsamples, static_cast<uint32_t>(a->nseries), a->reject_constant,
                         stats, flags, a->status);
      launch_check("stats_rowwise_kernel");
    }
    hipLaunchKernelGGL(normalise_kernel, g, dim3(kThreads), 0, st,
                       reinterpret_cast<float*>(data), a->stride, a->nsamples, stats, flags);
    launch_check("normalise_kernel");
  }
}

}  // namespace
}  // namespace gaffa_hip

namespace gaffa_hip {
namespace {

// Checks shared by both preprocess entry points; returns the scratch layout.
Layout validate(const gaffa_hip_preprocess_args* a) {
  check(a != nullptr, "preprocess arguments must not be null");
  check(a->data != nullptr && a->status != nullptr, "preprocess pointers must not be null");
  check(a->nseries > 0 && a->nsamples > 0, "preprocess batch must not be empty");
  check(a->stride >= a->nsamples, "preprocess stride must be >= nsamples");
  check(a->detrend || a->normalise || a->input_is_u32,
        "preprocess plan must contain at least one step");
  check(a->nseries < 65536, "preprocess batch has too many series for one launch");
  const uint64_t f = a->detrend ? a->scrunch_factor : 1;
  if (a->detrend) {
    check(a->median_window > 0 && a->median_window % 2 == 1,
          "running median detrend window must be odd and > 0");
    check(f >= 1, "scrunch factor must be >= 1");
    check(a->median_window + kThreads - 1 <= 4096,
          "GPU running median supports windows up to 3841 points");
    const uint64_t m = median_points(a->nsamples, f);
    if (f == 1) {
      check(a->median_window < a->nsamples,
            "running median detrend window_samples must be < input size");
    } else {
      check(m > a->median_window,
            "running median detrend low-resolution sample count must be > min_points");
    }
  }
  const Layout l = layout(a->nseries, a->nsamples, f);
  check_overflow(l.nblocks < 0x7FFFFFFFull, "preprocess series too long");
  return l;
}

// Scratch of the no-allocation entry point: the low-resolution series,
// medians, statistics and flags of one call.  Every kernel
// writes each element before reading it, so nothing needs clearing between
// calls.  Each device holds its own copy for the life of the process; calls
// that overlap on one device share it and race.
constexpr uint64_t kNoAllocScratchBytes = GAFFA_HIP_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` on the current device, looked up with the
// device current on the first call there and kept per device (host code
// never passes the symbol itself to a kernel).
void* noalloc_scratch() {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void* ptr[kMaxDevices] = {};
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices)
    throw status_error(GAFFA_HIP_RUNTIME, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (ptr[dev] == nullptr) {
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    if (p == nullptr) throw status_error(GAFFA_HIP_RUNTIME, "scratch symbol has no address");
    ptr[dev] = p;
  }
  return ptr[dev];
}

}  // namespace
}  // namespace gaffa_hip

using namespace gaffa_hip;

extern "C" {

uint64_t gaffa_hip_preprocess_workspace_bytes(uint64_t nseries, uint64_t nsamples,
                                              uint64_t scrunch_factor) {
  if (nseries == 0 || nsamples == 0 || scrunch_factor == 0) return 0;
  return layout(nseries, nsamples, scrunch_factor).total;
}

int gaffa_hip_detrend_normalise_batch(const gaffa_hip_preprocess_args* a,
                                      gaffa_hip_stream stream) {
  return guarded([&] {
    const Layout l = validate(a);
    check(a->workspace != nullptr && a->workspace_bytes >= l.total,
          "preprocess workspace is too small");
    if (a->input_is_u32) {
      run<uint32_t>(a, a->workspace, as_stream(stream));
    } else {
      run<float>(a, a->workspace, as_stream(stream));
    }
  });
}

int gaffa_hip_detrend_normalise_batch_noalloc(const gaffa_hip_preprocess_args* a,
                                              gaffa_hip_stream stream) {
  return guarded([&] {
    const Layout l = validate(a);
    check(a->workspace == nullptr && a->workspace_bytes == 0,
          "the no-allocation preprocess entry point uses its own scratch: workspace must be "
          "null and workspace_bytes 0");
    check_overflow(l.total <= kNoAllocScratchBytes,
                   "preprocess batch needs more scratch than the no-allocation entry point "
                   "holds (GAFFA_HIP_NOALLOC_SCRATCH_BYTES); split the batch");
    void* ws = noalloc_scratch();
    if (a->input_is_u32) {
      run<uint32_t>(a, ws, as_stream(stream));
    } else {
      run<float>(a, ws, as_stream(stream));
    }
  });
}

uint64_t gaffa_hip_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

}  // extern "C"
