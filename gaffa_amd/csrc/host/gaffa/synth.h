// Seeded synthetic filterbanks with injected pulsars (no gaffa counterpart;
// the benchmark and the parity tests generate their inputs with it).
#pragma once

#include "gaffa/gpu_memory.h"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace gaffa {

struct SynthPulsar {
  double period_s = 1.0;
  double dm = 0.0;
  double duty = 0.05;       // Gaussian FWHM as a fraction of the period
  double amp_sigma = 0.5;   // per-channel peak amplitude in noise sigma
  double phase0 = 0.0;      // pulse centre phase at t = 0
};

struct SynthFilterbankSpec {
  std::size_t nsamples = 0;
  std::size_t nchans = 0;
  double fch1 = 0.0;
  double foff = 0.0;
  double tsamp = 0.0;
  std::uint64_t seed = 0;
  std::vector<SynthPulsar> pulsars;
  std::vector<std::size_t> rfi_channels;
  std::int32_t rfi_offset = 0;
};

// Ascending channel frequencies, as gaffa's reader stores them
// (io/filterbank.cpp:79-118 reverses foff < 0 files).
std::vector<double> synth_frequencies(const SynthFilterbankSpec& spec);

struct SynthTables {
  std::vector<std::int32_t> delays;    // [pulsar][channel]
  std::vector<std::int32_t> profiles;  // [pulsar][4096], 1/4096-byte units
  std::vector<std::int64_t> period_q;  // samples * 2^20
  std::vector<std::int64_t> phase0_q;
  std::vector<std::int32_t> rfi_q;     // [channel]
  std::uint64_t key = 0;
};

SynthTables make_synth_tables(const SynthFilterbankSpec& spec);

// Generates samples [t_begin, t_begin + t_count) into device memory
// out[(t - t_begin) * nchans + c].
void synth_filterbank_gpu(const SynthFilterbankSpec& spec, std::uint8_t* out,
                          std::size_t t_begin, std::size_t t_count,
                          gaffa_hip_stream stream = nullptr);

}  // namespace gaffa
