// Learned candidate scorer (no gaffa counterpart): owns the network's
// parameters, Adam state and workspace on one device and drives the MFMA
// forward / backward kernels of libgaffa_hip (gaffa_hip_scorer_*).
//
// Inputs per candidate: a [32 subint][64 phase] folded cube (normalised to
// zero mean and unit variance) and a 128-trial periodogram patch (mean
// removed, S/N units), FP32.
#pragma once

#include "gaffa/gpu_memory.h"

#include <cstddef>
#include <cstdint>
#include <span>
#include <vector>

namespace gaffa {

inline constexpr std::size_t kScorerCubeFloats =
    GAFFA_HIP_SCORER_SUBINTS * GAFFA_HIP_SCORER_PHASE_BINS;
inline constexpr std::size_t kScorerPatchFloats = GAFFA_HIP_SCORER_PATCH;

struct ScorerTrainOptions {
  std::size_t epochs = 4;
  std::size_t batch = 256;
  float learning_rate = 1e-3f;
  float beta1 = 0.9f;
  float beta2 = 0.999f;
  float epsilon = 1e-8f;
  std::uint64_t seed = 1;  // minibatch order
};

class GpuCandidateScorer {
 public:
  explicit GpuCandidateScorer(std::size_t capacity = 1024, int device_id = 0);

  static gaffa_hip_scorer_layout layout();
  std::size_t parameter_count() const noexcept { return nparams_; }
  std::size_t capacity() const noexcept { return capacity_; }

  // He-uniform weights (bound sqrt(6 / fan_in)), zero biases, from a
  // splitmix64 stream of `seed`; resets Adam.
  void initialize(std::uint64_t seed);
  std::vector<float> parameters() const;
  void set_parameters(std::span<const float> values);

  // Logits of n candidates (host arrays [n][32*64] and [n][128]).
  std::vector<float> forward(const float* cubes, const float* patches, std::size_t n);
  // Mean BCE loss and every parameter gradient for one batch (n <= capacity).
  float gradients(const float* cubes, const float* patches, const float* labels, std::size_t n,
                  std::vector<float>& grads);
  // Adam training on a device-resident copy of the set; the loss of each
  // epoch (mean over its minibatches).
  std::vector<float> fit(const float* cubes, const float* patches, const float* labels,
                         std::size_t n, const ScorerTrainOptions& options);
  // Logits of n candidates already on this device.
  void forward_device(const float* cubes, const float* patches, std::size_t n, float* logits,
                      gaffa_hip_stream stream);

 private:
  gaffa_hip_scorer_net net() noexcept;
  std::size_t capacity_ = 0;
  std::size_t nparams_ = 0;
  int device_id_ = 0;
  GpuStream stream_;
  GpuDeviceBuffer<float> params_, grads_, adam_m_, adam_v_;
  GpuDeviceBuffer<std::uint16_t> params16_;
  GpuDeviceMemory workspace_;
  std::uint32_t step_ = 0;
};

}  // namespace gaffa
