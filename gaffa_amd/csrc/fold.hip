// Fractional-phase-overlap folding on gfx950 (gaffa's folding.cpp:300-420).
//
// Sample i covers the scaled phase interval [i * s, i * s + s), s = tsamp /
// period * nbin, and adds value * overlap to every bin it overlaps (bins wrap
// modulo nbin); exposure collects the overlaps; the folded value is
// sum / exposure where exposure > 0.
//
// One thread per (job, subint, phase bin, output channel) *gathers*: it walks
// the periods of its subint (bin k = phase + m * nbin for increasing m) and,
// per period, the few samples overlapping bin k in increasing time order.
// A sample overlaps at most one bin of a given phase (a period spans more than
// one sample), so this is exactly the order in which the CPU, walking samples
// in time order, adds to that bin: every FP64 sum equals the CPU's bit for bit
// and no atomics are needed.
#include "common.hpp"

#include <cmath>

namespace gaffa_hip {
namespace {

// The value of output channel c0 / factor at one sample.  `divisor` != 1
// marks an input whose elements are already sums of that many channels (exact
// integers), which the CPU's serial FP64 sum would reproduce exactly.
template <typename T>
__device__ __forceinline__ double sample_value(const T* row, uint64_t cstride, uint32_t c0,
                                               uint32_t factor, double divisor) {
  if (factor == 1) {
    const double v = static_cast<double>(row[c0 * cstride]);
    return divisor == 1.0 ? v : v / divisor;
  }
  double v = 0.0;  // serial in channel order, then the mean (folding.cpp:354-360)
  for (uint32_t o = 0; o < factor; ++o) v += static_cast<double>(row[(c0 + o) * cstride]);
  return v / static_cast<double>(factor);
}

template <typename T>
__global__ void __launch_bounds__(256)
fold_kernel(const T* __restrict__ data, gaffa_hip_fold_args a,
            const gaffa_hip_fold_job* __restrict__ jobs) {
  const gaffa_hip_fold_job job = jobs[blockIdx.y];
  const uint32_t nout = a.output_channels;
  const uint64_t per_job = static_cast<uint64_t>(a.nsubint) * nout * a.nbin;
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= per_job) return;
  // output channel fastest: the lanes of a wave read adjacent channel
  // groups of the same samples (whole rows per load) and walk the same
  // sample sequence (one phase per wave where nout >= 64)
  const uint32_t ch = static_cast<uint32_t>(e % nout);
  const uint32_t phase = static_cast<uint32_t>((e / nout) % a.nbin);
  const uint32_t sub = static_cast<uint32_t>(e / (static_cast<uint64_t>(a.nbin) * nout));
  const uint32_t factor = a.nchans / nout;

  const uint64_t t0 = static_cast<uint64_t>(sub) * a.subint_samples;
  const uint64_t t1 = min(a.nsamples, t0 + a.subint_samples);
  const double step = a.tsamp / job.period * static_cast<double>(a.nbin);
  const double nb = static_cast<double>(a.nbin);
  const int64_t i_first = static_cast<int64_t>(job.start_sample + t0);
  const int64_t i_last = static_cast<int64_t>(job.start_sample + t1) - 1;
  // bins k = phase + m * nbin that the subint's samples can reach
  const double lo_scaled = static_cast<double>(i_first) * step;
  const double hi_scaled = static_cast<double>(i_last) * step + step;
  const int64_t k_lo = static_cast<int64_t>(floor(lo_scaled));
  const int64_t k_hi = static_cast<int64_t>(ceil(hi_scaled) - 1.0);
  int64_t m = (k_lo - static_cast<int64_t>(phase)) / static_cast<int64_t>(a.nbin);
  if (static_cast<int64_t>(phase) + m * static_cast<int64_t>(a.nbin) < k_lo) ++m;
  const T* base = data + job.input_offset;
  const uint64_t cstride = a.channel_stride == 0 ? 1 : a.channel_stride;
  const double divisor = a.divisor == 0.0 ? 1.0 : a.divisor;

  double sum = 0.0, expo = 0.0;
  for (int64_t k = static_cast<int64_t>(phase) + m * static_cast<int64_t>(a.nbin); k <= k_hi;
       k += static_cast<int64_t>(a.nbin)) {
    const double kd = static_cast<double>(k);
    // samples with floor(i s) <= k <= ceil(i s + s) - 1, widened by one and
    // filtered by the CPU's own test below
    int64_t i = static_cast<int64_t>(floor(kd / step)) - 1;
    const int64_t i_end = static_cast<int64_t>(floor((kd + 1.0) / step)) + 1;
    if (i < i_first) i = i_first;
    for (; i <= i_end && i <= i_last; ++i) {
      const double low = static_cast<double>(i) * step;
      const double high = low + step;
      if (static_cast<int64_t>(floor(low)) > k || static_cast<int64_t>(ceil(high) - 1.0) < k)
        continue;
      const double w = fmin(high, kd + 1.0) - fmax(low, kd);
      if (!(w > 0.0)) continue;
      const uint64_t t = static_cast<uint64_t>(i) - job.start_sample;
      const double v = sample_value(base + t * a.row_stride, cstride, ch * factor, factor,
                                    divisor);
      sum += v * w;
      expo += w;
    }
  }
  const uint64_t out_index =
      (static_cast<uint64_t>(blockIdx.y) * a.nsubint * nout + static_cast<uint64_t>(sub) * nout +
       ch) * a.nbin + phase;
  a.out[out_index] = expo > 0.0 ? static_cast<float>(sum / expo) : 0.0f;
  if (ch == 0 && a.exposure != nullptr)
    a.exposure[(static_cast<uint64_t>(blockIdx.y) * a.nsubint + sub) * a.nbin + phase] = expo;
}

}  // namespace
}  // namespace gaffa_hip

using namespace gaffa_hip;

extern "C" int gaffa_hip_fold_batch(const gaffa_hip_fold_args* a, gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr && a->data != nullptr && a->jobs != nullptr && a->out != nullptr,
          "fold arguments must not be null");
    check(a->njobs > 0 && a->njobs < 65536, "fold batch must hold 1..65535 jobs");
    check(a->nbin > 0, "fold nbin must be positive");
    check(std::isfinite(a->tsamp) && a->tsamp > 0.0, "fold tsamp must be finite and positive");
    check(a->nsamples > 0 && a->nchans > 0, "fold spectrum shape must be non-empty");
    check(a->subint_samples > 0, "fold subint_samples must be positive");
    check(a->output_channels > 0 && a->output_channels <= a->nchans &&
              a->nchans % a->output_channels == 0,
          "fold input channel count must be divisible by output_channels");
    const uint64_t cstride = a->channel_stride == 0 ? 1 : a->channel_stride;
    check(cstride == 1 ? a->row_stride >= a->nchans
                       : a->row_stride == 1 && cstride >= a->nsamples,
          "fold strides must address every (sample, channel) once: sample-major rows of "
          "nchans, or channel-major rows (row_stride 1) of at least nsamples");
    check(a->divisor == 0.0 || (a->divisor >= 1.0 && std::isfinite(a->divisor)),
          "fold divisor must be 0 (none) or >= 1");
    check(a->nsubint == ceil_div(a->nsamples, a->subint_samples),
          "fold nsubint must equal ceil(nsamples / subint_samples)");
    const uint64_t per_job = a->nsubint * a->output_channels * a->nbin;
    check_overflow(per_job / 256 < (1ull << 31), "fold output too large");
    const dim3 grid(static_cast<uint32_t>(ceil_div(per_job, 256)),
                    static_cast<uint32_t>(a->njobs));
    const hipStream_t st = as_stream(stream);
    switch (a->dtype) {
      case GAFFA_HIP_FOLD_U8:
        hipLaunchKernelGGL(fold_kernel<uint8_t>, grid, dim3(256), 0, st,
                           static_cast<const uint8_t*>(a->data), *a, a->jobs);
        break;
      case GAFFA_HIP_FOLD_U16:
        hipLaunchKernelGGL(fold_kernel<uint16_t>, grid, dim3(256), 0, st,
                           static_cast<const uint16_t*>(a->data), *a, a->jobs);
        break;
      case GAFFA_HIP_FOLD_U32:
        hipLaunchKernelGGL(fold_kernel<uint32_t>, grid, dim3(256), 0, st,
                           static_cast<const uint32_t*>(a->data), *a, a->jobs);
        break;
      case GAFFA_HIP_FOLD_F32:
        hipLaunchKernelGGL(fold_kernel<float>, grid, dim3(256), 0, st,
                           static_cast<const float*>(a->data), *a, a->jobs);
        break;
      default:
        throw status_error(GAFFA_HIP_INVALID_ARGUMENT, "fold dtype must be u8, u16, u32 or f32");
    }
    launch_check("fold_kernel");
  });
}
