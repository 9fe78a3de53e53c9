// SIGPROC filterbank files: header parsing and a streamed reader that moves
// the samples to the device in chunks through pinned host buffers
// (include/gaffa/filterbank.h:28-64, src/gaffa/io/filterbank.cpp in gaffa).
// Channels are delivered in ascending frequency, as gaffa's reader stores
// them: a descending file (foff < 0 or a descending frequency table) is
// reversed on the device.
#pragma once

#include "gaffa/gpu_memory.h"

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

namespace gaffa {

// gaffa's read options (include/gaffa/filterbank.h:12-31): the delivered
// channel order, and how a descending file is reversed on the host.
enum class ChannelOrderPolicy { FrequencyAscending, PreserveFileOrder };
enum class ReverseBackend { Auto, CpuScalar, CpuOpenmp };
inline constexpr std::size_t default_filterbank_io_buffer_bytes = 64ull << 20;

struct FilterbankReadOptions {
  std::size_t io_buffer_bytes = default_filterbank_io_buffer_bytes;
  ChannelOrderPolicy channel_order = ChannelOrderPolicy::FrequencyAscending;
  ReverseBackend reverse_backend = ReverseBackend::Auto;
  std::size_t openmp_min_rows = 4096;
};

struct FilterbankHeader {
  std::size_t header_size = 0;
  std::size_t nsamples = 0;  // from the file size
  int telescope_id = 0, machine_id = 0, data_type = 1, barycentric = 0, pulsarcentric = 0;
  int ibeam = 0, nbeams = 0, npuls = 0, nbins = 0, nbits = 0, nifs = 1, nchans = 0;
  double az_start = 0.0, za_start = 0.0, src_raj = 0.0, src_dej = 0.0, tstart = 0.0;
  double tsamp = 0.0, fch1 = 0.0, foff = 0.0, refdm = 0.0, period = 0.0;
  std::string rawdatafile, source_name;
  std::vector<double> frequency_table;  // in the delivered channel order
  bool uses_frequency_table = false;
  bool reversed = false;  // channels are delivered in the reverse of the file's order
};

// Header of a SIGPROC file with the frequency metadata of the delivered
// channel order (ascending unless PreserveFileOrder), gaffa's checks and
// messages (io/filterbank.cpp:120-234).
FilterbankHeader read_filterbank_header(
    const std::string& path, ChannelOrderPolicy order = ChannelOrderPolicy::FrequencyAscending);

// Header and samples ([nsamples][nifs][nchans] elements of nbits / 8 bytes,
// little endian as stored) in the requested channel order
// (io/filterbank.cpp:236-373).
struct FilterbankData {
  FilterbankHeader header;
  std::vector<std::uint8_t> bytes;
};

FilterbankData read_filterbank(const std::string& path, const FilterbankReadOptions& options = {});

// Writes an 8-bit filterbank (channels as given, fch1/foff header).
void write_filterbank_u8(const std::string& path, const FilterbankHeader& header,
                         const std::uint8_t* samples, std::size_t nsamples);

struct FilterbankStreamStats {
  std::size_t bytes = 0;
  std::size_t chunks = 0;
  double read_ms = 0.0;   // host time in file reads
  double wall_ms = 0.0;   // whole stream, including the per-chunk work
};

// Streams samples [0, header.nsamples) of an 8-bit, nifs == 1 file into
// `device` ([nsamples][nchans], ascending channels) on `compute`, chunk by
// chunk: the file read of a chunk overlaps the copy and the device work of
// the previous ones.  After chunk k (samples [0, ready)) is on the device,
// on_chunk(ready) is called on the host to enqueue work that needs only
// those samples on `compute`.
FilterbankStreamStats stream_filterbank_u8(const std::string& path, const FilterbankHeader& header,
                                           std::uint8_t* device, gaffa_hip_stream compute,
                                           std::size_t chunk_samples,
                                           const std::function<void(std::size_t)>& on_chunk = {});

}  // namespace gaffa
