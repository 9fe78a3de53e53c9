// Folding (pfold): fractional-phase-overlap folds of dedispersed time series
// and dynamic spectra into profiles and subint x channel x phase cubes, with
// gaffa's types and results (include/gaffa/folding.h).  The folds run on the
// GPU through gaffa_hip_fold_batch; the small projections of a cube run on
// the host in gaffa's summation order.
#pragma once

#include "gaffa/gpu_memory.h"
#include "gaffa/sample_view.h"

#include <cstddef>
#include <cstdint>
#include <span>
#include <vector>

namespace gaffa {

struct FoldOptions {
  double period = 0.0;
  double tsamp = 0.0;
  std::size_t nbin = 0;
  std::size_t start_sample = 0;
};

struct FoldedProfile {
  std::vector<float> profile;
  std::vector<double> exposure;
  std::size_t nbin = 0;
};

struct FoldSpectrumOptions {
  double period = 0.0;
  double tsamp = 0.0;
  std::size_t nbin = 0;
  std::size_t subint_samples = 0;
  std::size_t start_sample = 0;
  std::size_t output_channels = 0;  // 0 keeps the input channel count
};

struct FoldedCube {
  std::vector<float> data;       // [nsubint][nchans][nbin]
  std::vector<double> exposure;  // [nsubint][nbin]
  std::size_t nsubint = 0;
  std::size_t nchans = 0;
  std::size_t nbin = 0;
};

struct FoldDedispersedSpectrumOptions {
  double period = 0.0;
  std::size_t nbin = 0;
  double tsubint = 10.0;
  std::size_t output_channels = 0;
};

struct FoldDedispersedSpectrumResult {
  FoldedCube cube;
  std::vector<float> profile;     // [nbin]
  std::vector<float> freq_phase;  // [nchans][nbin]
  std::vector<float> time_phase;  // [nsubint][nbin]
  std::size_t nsubint = 0;
  std::size_t nchans = 0;
  std::size_t nbin = 0;
  double period = 0.0;
  double tsamp = 0.0;
  double tsubint = 0.0;
};

// Device inputs: samples [nsamples][nchans] (row_stride elements apart).
enum class FoldDtype { U8 = 0, U16 = 1, U32 = 2, F32 = 3 };

struct GpuFoldInput {
  const void* data = nullptr;
  FoldDtype dtype = FoldDtype::F32;
  std::size_t nsamples = 0;
  std::size_t nchans = 1;
  std::size_t row_stride = 0;  // 0 = nchans
  int device_id = 0;
  std::size_t channel_stride = 0;  // elements between channels, 0 = 1 (sample-major rows)
  double divisor = 0.0;            // elements are sums of this many channels, 0 = 1
};

// One fold of a batch: input rows start at data + input_offset elements.
struct FoldJob {
  std::size_t input_offset = 0;
  double period = 0.0;
  std::size_t start_sample = 0;
};

// Profile (one subint over every sample).  gaffa sums with an OpenMP
// reduction whose order follows its thread count; this sums in time order.
FoldedProfile fold_time_series_gpu(const GpuFoldInput& input, const FoldOptions& options,
                                   gaffa_hip_stream stream = nullptr);

// Subint cube, equal to fold_spectrum_cpu bit for bit (period > tsamp).
FoldedCube fold_spectrum_gpu(const GpuFoldInput& input, const FoldSpectrumOptions& options,
                             gaffa_hip_stream stream = nullptr);

// Cube + profile / frequency-phase / time-phase projections
// (folding.cpp:237-297, 405-430).
FoldDedispersedSpectrumResult fold_dedispersed_spectrum_gpu(
    const GpuFoldInput& input, double tsamp, const FoldDedispersedSpectrumOptions& options,
    gaffa_hip_stream stream = nullptr);

// Many folds of one input layout in one launch (e.g. candidates of a DM x
// time plane, each its own row and period): cubes [njobs][nsubint][nchans_out]
// [nbin] and exposure [njobs][nsubint][nbin].
// As fold_batch_gpu, leaving the folded values on the device in `out`
// ([jobs][nsubint][output_channels][nbin] floats, enqueued on `stream`).
void fold_batch_gpu_device(const GpuFoldInput& input, std::span<const FoldJob> jobs, double tsamp,
                           std::size_t nbin, std::size_t subint_samples,
                           std::size_t output_channels, float* out, gaffa_hip_stream stream);

FoldedCube fold_batch_gpu(const GpuFoldInput& input, std::span<const FoldJob> jobs, double tsamp,
                          std::size_t nbin, std::size_t subint_samples,
                          std::size_t output_channels, std::vector<double>* exposure,
                          gaffa_hip_stream stream = nullptr);

FoldDedispersedSpectrumResult project_folded_cube(FoldedCube cube,
                                                  const FoldDedispersedSpectrumOptions& options,
                                                  double tsamp, double actual_tsubint);

std::size_t fold_subint_samples(double tsubint, double tsamp);

}  // namespace gaffa
