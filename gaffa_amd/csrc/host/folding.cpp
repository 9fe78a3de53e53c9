// Folding on the GPU (see gaffa/folding.h).
#include "gaffa/folding.h"

#include <cmath>
#include <limits>
#include <stdexcept>

namespace gaffa {
namespace {

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

void validate_shape(double period, double tsamp, std::size_t nbin) {
  check_arg(std::isfinite(period) && period > 0.0, "fold period must be finite and positive");
  check_arg(std::isfinite(tsamp) && tsamp > 0.0, "fold tsamp must be finite and positive");
  check_arg(nbin != 0, "fold nbin must be positive");
  if (nbin > 0xFFFFFFFFull) throw std::overflow_error("fold nbin exceeds the GPU range");
  // a sample spanning a whole period would add to one phase bin twice out of
  // time order; gaffa's folds are only meaningful for period > tsamp
  check_arg(period > tsamp, "fold period must exceed tsamp");
}

void validate_phase_range(std::size_t start, std::size_t count, double step) {
  const double hi = static_cast<double>(start + count - 1) * step + step;
  if (!std::isfinite(hi) || hi >= 9.0e18) {
    throw std::overflow_error("fold phase bin index exceeds int64 range");
  }
}

}  // namespace

std::size_t fold_subint_samples(double tsubint, double tsamp) {
  check_arg(std::isfinite(tsubint) && tsubint > 0.0, "fold tsubint must be finite and positive");
  check_arg(std::isfinite(tsamp) && tsamp > 0.0, "fold spectrum tsamp must be finite and positive");
  const double rounded = std::round(tsubint / tsamp);
  check_arg(std::isfinite(rounded) && rounded < 1.8e19, "fold subint sample count is too large");
  return std::max<std::size_t>(1, static_cast<std::size_t>(rounded));
}

namespace {

// Validates, uploads the jobs and launches the fold kernel into out /
// exposure (device).  Returns nsubint.
std::size_t launch_fold(const GpuFoldInput& in, std::span<const FoldJob> jobs, double tsamp,
                        std::size_t nbin, std::size_t subint_samples, std::size_t nout,
                        float* out, double* exposure, gaffa_hip_stream stream) {
  const std::size_t nsubint = (in.nsamples + subint_samples - 1) / subint_samples;
  std::vector<gaffa_hip_fold_job> hj;
  for (const FoldJob& j : jobs) hj.push_back(gaffa_hip_fold_job{j.input_offset, j.period, j.start_sample});
  GpuDeviceBuffer<gaffa_hip_fold_job> d_jobs = upload(hj.data(), hj.size());
  gaffa_hip_fold_args a{};
  a.data = in.data;
  a.dtype = static_cast<std::int32_t>(in.dtype);
  a.nsamples = in.nsamples;
  a.nchans = static_cast<std::uint32_t>(in.nchans);
  a.row_stride = in.row_stride == 0 ? in.nchans : in.row_stride;
  a.tsamp = tsamp;
  a.nbin = static_cast<std::uint32_t>(nbin);
  a.subint_samples = subint_samples;
  a.nsubint = nsubint;
  a.output_channels = static_cast<std::uint32_t>(nout);
  a.jobs = d_jobs.data();
  a.njobs = static_cast<std::uint32_t>(jobs.size());
  a.out = out;
  a.exposure = exposure;
  a.channel_stride = in.channel_stride;
  a.divisor = in.divisor;
  hip_ok(gaffa_hip_fold_batch(&a, stream), "fold_batch");
  // the job table must outlive the kernel
  hip_ok(gaffa_hip_stream_synchronize(stream), "hipStreamSynchronize");
  return nsubint;
}

}  // namespace

void fold_batch_gpu_device(const GpuFoldInput& in, std::span<const FoldJob> jobs, double tsamp,
                           std::size_t nbin, std::size_t subint_samples,
                           std::size_t output_channels, float* out, gaffa_hip_stream stream) {
  check_arg(in.data != nullptr && out != nullptr, "fold data and output must not be null");
  check_arg(in.nsamples != 0 && in.nchans != 0, "fold spectrum shape must be non-empty");
  check_arg(!jobs.empty(), "fold batch must contain at least one job");
  check_arg(subint_samples != 0, "fold subint_samples must be positive");
  const std::size_t nout = output_channels == 0 ? in.nchans : output_channels;
  check_arg(nout <= in.nchans && in.nchans % nout == 0,
            "fold output_channels must divide the input channel count");
  for (const FoldJob& j : jobs) {
    validate_shape(j.period, tsamp, nbin);
    validate_phase_range(j.start_sample, in.nsamples, tsamp / j.period * static_cast<double>(nbin));
  }
  GpuDeviceScope scope(in.device_id);
  const std::size_t nsubint = (in.nsamples + subint_samples - 1) / subint_samples;
  GpuDeviceBuffer<double> d_exp(jobs.size() * nsubint * nbin);
  launch_fold(in, jobs, tsamp, nbin, subint_samples, nout, out, d_exp.data(), stream);
}

FoldedCube fold_batch_gpu(const GpuFoldInput& in, std::span<const FoldJob> jobs, double tsamp,
                          std::size_t nbin, std::size_t subint_samples,
                          std::size_t output_channels, std::vector<double>* exposure,
                          gaffa_hip_stream stream) {
  check_arg(in.data != nullptr, "fold spectrum data must not be null");
  check_arg(in.nsamples != 0 && in.nchans != 0, "fold spectrum shape must be non-empty");
  check_arg(!jobs.empty(), "fold batch must contain at least one job");
  check_arg(subint_samples != 0, "fold subint_samples must be positive");
  const std::size_t nout = output_channels == 0 ? in.nchans : output_channels;
  check_arg(nout <= in.nchans, "fold output_channels must be <= input channel count");
  check_arg(in.nchans % nout == 0, "fold input channel count must be divisible by output_channels");
  for (const FoldJob& j : jobs) {
    validate_shape(j.period, tsamp, nbin);
    validate_phase_range(j.start_sample, in.nsamples, tsamp / j.period * static_cast<double>(nbin));
  }
  const std::size_t nsubint = (in.nsamples + subint_samples - 1) / subint_samples;
  GpuDeviceScope scope(in.device_id);
  std::vector<gaffa_hip_fold_job> hj;
  for (const FoldJob& j : jobs) hj.push_back(gaffa_hip_fold_job{j.input_offset, j.period, j.start_sample});
  GpuDeviceBuffer<gaffa_hip_fold_job> d_jobs = upload(hj.data(), hj.size());
  const std::size_t cube = nsubint * nout * nbin;
  GpuDeviceBuffer<float> d_out(jobs.size() * cube);
  GpuDeviceBuffer<double> d_exp(jobs.size() * nsubint * nbin);
  gaffa_hip_fold_args a{};
  a.data = in.data;
  a.dtype = static_cast<std::int32_t>(in.dtype);
  a.nsamples = in.nsamples;
  a.nchans = static_cast<std::uint32_t>(in.nchans);
  a.row_stride = in.row_stride == 0 ? in.nchans : in.row_stride;
  a.tsamp = tsamp;
  a.nbin = static_cast<std::uint32_t>(nbin);
  a.subint_samples = subint_samples;
  a.nsubint = nsubint;
  a.output_channels = static_cast<std::uint32_t>(nout);
  a.jobs = d_jobs.data();
  a.njobs = static_cast<std::uint32_t>(jobs.size());
  a.out = d_out.data();
  a.exposure = d_exp.data();
  hip_ok(gaffa_hip_fold_batch(&a, stream), "fold_batch");
  FoldedCube r;
  r.data.resize(d_out.size());
  r.exposure.resize(d_exp.size());
  copy_to_host(r.data.data(), d_out.data(), d_out.bytes(), stream);
  copy_to_host(r.exposure.data(), d_exp.data(), d_exp.bytes(), stream);
  r.nsubint = nsubint;
  r.nchans = nout;
  r.nbin = nbin;
  if (exposure != nullptr) *exposure = r.exposure;
  return r;
}

FoldedProfile fold_time_series_gpu(const GpuFoldInput& in, const FoldOptions& o,
                                   gaffa_hip_stream stream) {
  validate_shape(o.period, o.tsamp, o.nbin);
  check_arg(in.nsamples != 0, "fold input must not be empty");
  const FoldJob job{0, o.period, o.start_sample};
  FoldedCube c = fold_batch_gpu(in, std::span<const FoldJob>(&job, 1), o.tsamp, o.nbin,
                                in.nsamples, 1, nullptr, stream);
  return FoldedProfile{std::move(c.data), std::move(c.exposure), o.nbin};
}

FoldedCube fold_spectrum_gpu(const GpuFoldInput& in, const FoldSpectrumOptions& o,
                             gaffa_hip_stream stream) {
  validate_shape(o.period, o.tsamp, o.nbin);
  const FoldJob job{0, o.period, o.start_sample};
  return fold_batch_gpu(in, std::span<const FoldJob>(&job, 1), o.tsamp, o.nbin, o.subint_samples,
                        o.output_channels, nullptr, stream);
}

FoldDedispersedSpectrumResult project_folded_cube(FoldedCube cube,
                                                  const FoldDedispersedSpectrumOptions& o,
                                                  double tsamp, double actual_tsubint) {
  // Exposure-weighted means over subints (profile, freq_phase) and the
  // channel mean per subint (time_phase), summed subint by subint, phase by
  // phase, channel by channel as gaffa does (folding.cpp:237-297).
  const std::size_t ns = cube.nsubint, nc = cube.nchans, nb = cube.nbin;
  FoldDedispersedSpectrumResult r;
  r.profile.assign(nb, 0.0F);
  r.freq_phase.assign(nc * nb, 0.0F);
  r.time_phase.assign(ns * nb, 0.0F);
  std::vector<double> psum(nb, 0.0), pexp(nb, 0.0), fsum(nc * nb, 0.0), fexp(nc * nb, 0.0);
  for (std::size_t s = 0; s < ns; ++s) {
    for (std::size_t b = 0; b < nb; ++b) {
      const double e = cube.exposure[s * nb + b];
      if (!(e > 0.0)) continue;
      double tsum = 0.0;
      for (std::size_t c = 0; c < nc; ++c) {
        const float v = cube.data[(s * nc + c) * nb + b];
        tsum += v;
        psum[b] += static_cast<double>(v) * e;
        pexp[b] += e;
        fsum[c * nb + b] += static_cast<double>(v) * e;
        fexp[c * nb + b] += e;
      }
      r.time_phase[s * nb + b] = static_cast<float>(tsum / static_cast<double>(nc));
    }
  }
  for (std::size_t b = 0; b < nb; ++b)
    if (pexp[b] > 0.0) r.profile[b] = static_cast<float>(psum[b] / pexp[b]);
  for (std::size_t i = 0; i < fsum.size(); ++i)
    if (fexp[i] > 0.0) r.freq_phase[i] = static_cast<float>(fsum[i] / fexp[i]);
  r.nsubint = ns;
  r.nchans = nc;
  r.nbin = nb;
  r.period = o.period;
  r.tsamp = tsamp;
  r.tsubint = actual_tsubint;
  r.cube = std::move(cube);
  return r;
}

FoldDedispersedSpectrumResult fold_dedispersed_spectrum_gpu(
    const GpuFoldInput& in, double tsamp, const FoldDedispersedSpectrumOptions& o,
    gaffa_hip_stream stream) {
  const std::size_t sub = fold_subint_samples(o.tsubint, tsamp);
  FoldedCube cube = fold_spectrum_gpu(
      in, FoldSpectrumOptions{o.period, tsamp, o.nbin, sub, 0, o.output_channels}, stream);
  return project_folded_cube(std::move(cube), o, tsamp, static_cast<double>(sub) * tsamp);
}

}  // namespace gaffa
