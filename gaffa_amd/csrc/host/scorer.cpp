// Host side of the candidate scorer: parameter initialisation, batching and
// the training loop around the gaffa_hip_scorer_* kernels.
#include "gaffa/scorer.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <random>
#include <stdexcept>

namespace gaffa {
namespace {

std::uint64_t splitmix64(std::uint64_t& s) {
  std::uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

void check_arg(bool ok, const char* message) {
  if (!ok) throw std::invalid_argument(message);
}

}  // namespace

gaffa_hip_scorer_layout GpuCandidateScorer::layout() {
  gaffa_hip_scorer_layout l{};
  hip_ok(gaffa_hip_scorer_get_layout(&l), "scorer_get_layout");
  return l;
}

GpuCandidateScorer::GpuCandidateScorer(std::size_t capacity, int device_id)
    : capacity_(capacity), device_id_(device_id) {
  check_arg(capacity > 0 && capacity <= 65535, "scorer capacity must be in 1..65535");
  GpuDeviceScope scope(device_id);
  nparams_ = static_cast<std::size_t>(layout().nparams);
  stream_ = GpuStream::create();
  params_ = GpuDeviceBuffer<float>(nparams_);
  grads_ = GpuDeviceBuffer<float>(nparams_);
  adam_m_ = GpuDeviceBuffer<float>(nparams_);
  adam_v_ = GpuDeviceBuffer<float>(nparams_);
  params16_ = GpuDeviceBuffer<std::uint16_t>(nparams_);
  workspace_ = GpuDeviceMemory(
      static_cast<std::size_t>(gaffa_hip_scorer_workspace_bytes(static_cast<std::uint32_t>(capacity))));
  initialize(1);
}

gaffa_hip_scorer_net GpuCandidateScorer::net() noexcept {
  gaffa_hip_scorer_net n{};
  n.params = params_.data();
  n.params16 = params16_.data();
  n.grads = grads_.data();
  n.adam_m = adam_m_.data();
  n.adam_v = adam_v_.data();
  n.workspace = workspace_.data();
  n.capacity = static_cast<std::uint32_t>(capacity_);
  return n;
}

void GpuCandidateScorer::initialize(std::uint64_t seed) {
  const gaffa_hip_scorer_layout l = layout();
  std::vector<float> p(nparams_, 0.0f);
  std::uint64_t s = seed;
  for (int layer = 0; layer < GAFFA_HIP_SCORER_LAYERS; ++layer) {
    const double bound = std::sqrt(6.0 / static_cast<double>(l.fan_in[layer]));
    for (std::uint64_t i = l.weight[layer]; i < l.bias[layer]; ++i) {
      const double u = static_cast<double>(splitmix64(s) >> 11) * 0x1.0p-53;  // [0, 1)
      p[i] = static_cast<float>((2.0 * u - 1.0) * bound);
    }
  }
  set_parameters(p);
}

std::vector<float> GpuCandidateScorer::parameters() const {
  GpuDeviceScope scope(device_id_);
  std::vector<float> p(nparams_);
  copy_to_host(p.data(), params_.data(), nparams_ * sizeof(float), stream_.get());
  stream_.synchronize();
  return p;
}

void GpuCandidateScorer::set_parameters(std::span<const float> values) {
  check_arg(values.size() == nparams_, "scorer parameter vector has the wrong length");
  for (float v : values) check_arg(std::isfinite(v), "scorer parameters must be finite");
  GpuDeviceScope scope(device_id_);
  const gaffa_hip_scorer_net n = net();
  copy_to_device(params_.data(), values.data(), nparams_ * sizeof(float), stream_.get());
  hip_ok(gaffa_hip_fill_u32(reinterpret_cast<std::uint32_t*>(adam_m_.data()), 0, nparams_,
                            stream_.get()),
         "fill_u32");
  hip_ok(gaffa_hip_fill_u32(reinterpret_cast<std::uint32_t*>(adam_v_.data()), 0, nparams_,
                            stream_.get()),
         "fill_u32");
  hip_ok(gaffa_hip_scorer_sync_params(&n, stream_.get()), "scorer_sync_params");
  stream_.synchronize();
  step_ = 0;
}

void GpuCandidateScorer::forward_device(const float* cubes, const float* patches, std::size_t n,
                                        float* logits, gaffa_hip_stream stream) {
  GpuDeviceScope scope(device_id_);
  const gaffa_hip_scorer_net nt = net();
  for (std::size_t b0 = 0; b0 < n; b0 += capacity_) {
    const std::size_t b = std::min(capacity_, n - b0);
    hip_ok(gaffa_hip_scorer_forward(&nt, cubes + b0 * kScorerCubeFloats,
                                    patches + b0 * kScorerPatchFloats, nullptr,
                                    static_cast<std::uint32_t>(b), logits + b0, stream),
           "scorer_forward");
  }
}

std::vector<float> GpuCandidateScorer::forward(const float* cubes, const float* patches,
                                               std::size_t n) {
  check_arg(n > 0 && cubes != nullptr && patches != nullptr, "scorer input is empty");
  GpuDeviceScope scope(device_id_);
  GpuDeviceBuffer<float> d_cubes = upload(cubes, n * kScorerCubeFloats);
  GpuDeviceBuffer<float> d_patches = upload(patches, n * kScorerPatchFloats);
  GpuDeviceBuffer<float> d_logits(n);
  forward_device(d_cubes.data(), d_patches.data(), n, d_logits.data(), stream_.get());
  std::vector<float> out(n);
  copy_to_host(out.data(), d_logits.data(), n * sizeof(float), stream_.get());
  stream_.synchronize();
  return out;
}

float GpuCandidateScorer::gradients(const float* cubes, const float* patches, const float* labels,
                                    std::size_t n, std::vector<float>& grads) {
  check_arg(n > 0 && n <= capacity_, "scorer gradient batch must be in 1..capacity");
  GpuDeviceScope scope(device_id_);
  GpuDeviceBuffer<float> d_cubes = upload(cubes, n * kScorerCubeFloats);
  GpuDeviceBuffer<float> d_patches = upload(patches, n * kScorerPatchFloats);
  GpuDeviceBuffer<float> d_labels = upload(labels, n);
  GpuDeviceBuffer<float> d_logits(n), d_loss(1);
  const gaffa_hip_scorer_net nt = net();
  const auto b = static_cast<std::uint32_t>(n);
  hip_ok(gaffa_hip_scorer_forward(&nt, d_cubes.data(), d_patches.data(), nullptr, b,
                                  d_logits.data(), stream_.get()),
         "scorer_forward");
  hip_ok(gaffa_hip_scorer_backward(&nt, d_labels.data(), nullptr, b, d_loss.data(),
                                   stream_.get()),
         "scorer_backward");
  grads.resize(nparams_);
  float loss = 0.0f;
  copy_to_host(grads.data(), grads_.data(), nparams_ * sizeof(float), stream_.get());
  copy_to_host(&loss, d_loss.data(), sizeof(float), stream_.get());
  stream_.synchronize();
  return loss;
}

std::vector<float> GpuCandidateScorer::fit(const float* cubes, const float* patches,
                                           const float* labels, std::size_t n,
                                           const ScorerTrainOptions& o) {
  check_arg(n > 0, "scorer training set is empty");
  check_arg(o.batch > 0 && o.batch <= capacity_, "scorer batch must be in 1..capacity");
  check_arg(o.epochs > 0, "scorer epochs must be positive");
  for (std::size_t i = 0; i < n; ++i)
    check_arg(labels[i] == 0.0f || labels[i] == 1.0f, "scorer labels must be 0 or 1");
  GpuDeviceScope scope(device_id_);
  GpuDeviceBuffer<float> d_cubes = upload(cubes, n * kScorerCubeFloats);
  GpuDeviceBuffer<float> d_patches = upload(patches, n * kScorerPatchFloats);
  GpuDeviceBuffer<float> d_labels = upload(labels, n);
  const std::size_t steps = n / o.batch;  // whole minibatches only
  check_arg(steps > 0, "scorer training set is smaller than one batch");
  GpuDeviceBuffer<std::uint32_t> d_index(n);
  GpuDeviceBuffer<float> d_logits(o.batch), d_loss(steps);
  const gaffa_hip_scorer_net nt = net();
  std::vector<std::uint32_t> order(n);
  std::mt19937_64 rng(o.seed);
  std::vector<float> epoch_loss;
  std::vector<float> losses(steps);
  for (std::size_t e = 0; e < o.epochs; ++e) {
    std::iota(order.begin(), order.end(), 0u);
    std::shuffle(order.begin(), order.end(), rng);
    copy_to_device(d_index.data(), order.data(), n * sizeof(std::uint32_t), stream_.get());
    for (std::size_t s = 0; s < steps; ++s) {
      const std::uint32_t* idx = d_index.data() + s * o.batch;
      const auto b = static_cast<std::uint32_t>(o.batch);
      hip_ok(gaffa_hip_scorer_forward(&nt, d_cubes.data(), d_patches.data(), idx, b,
                                      d_logits.data(), stream_.get()),
             "scorer_forward");
      hip_ok(gaffa_hip_scorer_backward(&nt, d_labels.data(), idx, b, d_loss.data() + s,
                                       stream_.get()),
             "scorer_backward");
      hip_ok(gaffa_hip_scorer_adam(&nt, o.learning_rate, o.beta1, o.beta2, o.epsilon, ++step_,
                                   stream_.get()),
             "scorer_adam");
    }
    copy_to_host(losses.data(), d_loss.data(), steps * sizeof(float), stream_.get());
    stream_.synchronize();
    double sum = 0.0;
    for (float l : losses) sum += l;
    epoch_loss.push_back(static_cast<float>(sum / static_cast<double>(steps)));
  }
  return epoch_loss;
}

}  // namespace gaffa
