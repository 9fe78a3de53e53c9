// Seeded synthetic filterbank generator (no gaffa counterpart; it feeds the
// benchmark and the tests).  Sample (t, c) of a [t][c] uint8 buffer is
//   clip(pop - 384 + floor((dither + rfi_q[c] + sum_p profile_p[bin_p]) / 4096))
// where pop is the popcount of 16 hashed 64-bit words (Binomial(1024, 1/2):
// mean 512, sigma 16), dither is 12 hashed bits, and bin_p is the exact
// integer pulse phase ((t - delay_p[c]) * 2^20 - phase0_q) mod period_q,
// scaled to 4096 bins.  Only integer arithmetic runs on the device, so the
// bytes equal the CPU oracle's for the same host tables.
#include "common.hpp"

#include <cmath>

namespace gaffa_hip {
namespace {

constexpr int kPhaseBins = 4096;
constexpr int64_t kPeriodFrac = 1LL << 20;

__device__ __forceinline__ uint64_t splitmix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(256)
synth_kernel(gaffa_hip_synth_args a) {
  const uint64_t total = a.t_count * a.nchans;
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= total) return;
  const uint64_t tl = e / a.nchans;
  const uint64_t c = e - tl * a.nchans;
  const uint64_t t = a.t_begin + tl;
  const uint64_t k = (t * a.nchans + c) * 17;
  int64_t pop = 0;
#pragma unroll
  for (uint64_t j = 0; j < 16; ++j) pop += __popcll(splitmix(a.key + k + j));
  int64_t acc = static_cast<int64_t>(splitmix(a.key + k + 16) & 4095u);
  acc += a.rfi_q[c];
  for (uint32_t p = 0; p < a.npulsars; ++p) {
    const int64_t tau = static_cast<int64_t>(t) - a.delays[p * a.nchans + c];
    const int64_t pq = a.period_q[p];
    int64_t ph = (tau * kPeriodFrac - a.phase0_q[p]) % pq;
    if (ph < 0) ph += pq;
    const int64_t bin = ph * kPhaseBins / pq;
    acc += a.profiles[static_cast<uint64_t>(p) * kPhaseBins + bin];
  }
  const int64_t q = acc >= 0 ? acc / 4096 : -((-acc + 4095) / 4096);
  int64_t v = pop - 384 + q;
  v = v < 0 ? 0 : (v > 255 ? 255 : v);
  a.out[e] = static_cast<uint8_t>(v);
}

// One sample of a packed batch of float series: N(0,1) from two hashed
// uniforms (Box-Muller), plus a Gaussian pulse train for injected series.
__global__ void __launch_bounds__(256)
synth_series_kernel(gaffa_hip_synth_series_args a, uint64_t e0) {
  const uint64_t e = e0 + static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= a.row_length) return;
  // series s owns [offsets[s], offsets[s] + lengths[s]); the tail of its
  // 64-float padding is written as 0
  uint32_t lo = 0, hi = a.nseries;  // last series with offsets[s] <= e
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (a.offsets[mid] <= e) lo = mid; else hi = mid;
  }
  const uint64_t t = e - a.offsets[lo];
  if (t >= a.lengths[lo]) {
    a.out[e] = 0.0f;
    return;
  }
  const uint64_t h = splitmix(a.key ^ (static_cast<uint64_t>(lo) << 40) ^ t);
  const float u1 = (static_cast<float>(h >> 40) + 1.0f) * (1.0f / 16777217.0f);  // (0, 1]
  const float u2 = static_cast<float>((h >> 16) & 0xFFFFFFu) * (1.0f / 16777216.0f);
  float v = sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
  const float amp = a.amplitude[lo];
  if (amp != 0.0f) {
    const double ph = fmod(static_cast<double>(t) * a.tsamp / a.period[lo] + a.phase0[lo], 1.0);
    const float d = static_cast<float>(ph - 0.5) / a.sigma[lo];
    v += amp * expf(-0.5f * d * d);
  }
  a.out[e] = v;
}

}  // namespace
}  // namespace gaffa_hip

using namespace gaffa_hip;

extern "C" int gaffa_hip_synth_series(const gaffa_hip_synth_series_args* a,
                                      gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr && a->out != nullptr && a->offsets != nullptr && a->lengths != nullptr &&
              a->period != nullptr && a->phase0 != nullptr && a->sigma != nullptr &&
              a->amplitude != nullptr,
          "synth arguments must not be null");
    check(a->nseries > 0 && a->tsamp > 0.0, "synth batch must be non-empty with tsamp > 0");
    for (uint64_t e0 = 0; e0 < a->row_length; e0 += (1ull << 30)) {
      const uint64_t count = std::min<uint64_t>(1ull << 30, a->row_length - e0);
      hipLaunchKernelGGL(synth_series_kernel, dim3(static_cast<uint32_t>(ceil_div(count, 256))),
                         dim3(256), 0, as_stream(stream), *a, e0);
      launch_check("synth_series_kernel");
    }
  });
}

extern "C" int gaffa_hip_synth_filterbank(const gaffa_hip_synth_args* a,
                                          gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr && a->out != nullptr && a->rfi_q != nullptr,
          "synth arguments must not be null");
    check(a->nchans > 0, "synth requires at least one channel");
    check(a->npulsars == 0 ||
              (a->delays != nullptr && a->profiles != nullptr && a->period_q != nullptr &&
               a->phase0_q != nullptr),
          "synth pulsar tables must not be null");
    if (a->t_count * a->nchans == 0) return;
    // one launch covers at most 2^30 samples (the grid's total work-item
    // count must stay below 2^32)
    const uint64_t rows_per_launch = std::max<uint64_t>(1, (1ull << 30) / a->nchans);
    for (uint64_t t0 = 0; t0 < a->t_count; t0 += rows_per_launch) {
      gaffa_hip_synth_args part = *a;
      part.t_begin = a->t_begin + t0;
      part.t_count = std::min(rows_per_launch, a->t_count - t0);
      part.out = a->out + t0 * a->nchans;
      const uint64_t blocks = ceil_div(part.t_count * part.nchans, 256);
      check_overflow(blocks <= (1ull << 22), "synth launch too large");
      hipLaunchKernelGGL(synth_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                         as_stream(stream), part);
      launch_check("synth_kernel");
    }
  });
}
