// SIGPROC filterbank header parsing and streamed device upload.
#include "gaffa/filterbank.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cmath>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <set>
#include <stdexcept>

namespace gaffa {
namespace {

struct File {
  std::FILE* f = nullptr;
  explicit File(const std::string& path, const char* mode) : f(std::fopen(path.c_str(), mode)) {
    if (f == nullptr) throw std::runtime_error("cannot open filterbank file: " + path);
  }
  ~File() {
    if (f != nullptr) std::fclose(f);
  }
  File(const File&) = delete;
  File& operator=(const File&) = delete;
};

template <typename T>
T read_value(std::FILE* f, const char* field) {
  T v{};
  if (std::fread(&v, sizeof(T), 1, f) != 1)
    throw std::runtime_error(std::string("Failed to read filterbank field: ") + field);
  return v;
}

std::string read_string(std::FILE* f) {
  const std::int32_t n = read_value<std::int32_t>(f, "string length");
  if (n <= 0 || n > 4096) throw std::runtime_error("Invalid filterbank string length");
  std::string s(static_cast<std::size_t>(n), '\0');
  if (std::fread(s.data(), 1, s.size(), f) != s.size())
    throw std::runtime_error("Failed to read filterbank string");
  return s;
}

void write_string(std::FILE* f, const std::string& s) {
  const std::int32_t n = static_cast<std::int32_t>(s.size());
  std::fwrite(&n, 4, 1, f);
  std::fwrite(s.data(), 1, s.size(), f);
}

template <typename T>
void write_field(std::FILE* f, const char* key, T v) {
  write_string(f, key);
  std::fwrite(&v, sizeof(T), 1, f);
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace

namespace {

// Parses the header from the current position of `f` (the file start) and
// sizes the sample block from the file length.  Keyword handling and checks
// follow gaffa's reader (io/filterbank.cpp:120-234): strings take the next
// string, FREQUENCY_START ... FREQUENCY_END brackets the fchannel table (an
// fchannel outside it is an error; a table zeroes fch1 / foff), nsamples in
// the header is ignored, and the data must be whole rows.
FilterbankHeader parse_header(std::FILE* f, ChannelOrderPolicy order) {
  FilterbankHeader h;
  if (read_string(f) != "HEADER_START")
    throw std::runtime_error("Non-standard filterbank file: missing HEADER_START");
  struct IntKey {
    const char* key;
    int FilterbankHeader::*field;
  };
  struct DoubleKey {
    const char* key;
    double FilterbankHeader::*field;
  };
  static const IntKey ints[] = {
      {"nchans", &FilterbankHeader::nchans},         {"telescope_id", &FilterbankHeader::telescope_id},
      {"machine_id", &FilterbankHeader::machine_id}, {"data_type", &FilterbankHeader::data_type},
      {"ibeam", &FilterbankHeader::ibeam},           {"nbeams", &FilterbankHeader::nbeams},
      {"nbits", &FilterbankHeader::nbits},           {"barycentric", &FilterbankHeader::barycentric},
      {"pulsarcentric", &FilterbankHeader::pulsarcentric},
      {"nbins", &FilterbankHeader::nbins},           {"nifs", &FilterbankHeader::nifs},
      {"npuls", &FilterbankHeader::npuls}};
  static const DoubleKey doubles[] = {
      {"az_start", &FilterbankHeader::az_start}, {"za_start", &FilterbankHeader::za_start},
      {"src_raj", &FilterbankHeader::src_raj},   {"src_dej", &FilterbankHeader::src_dej},
      {"tstart", &FilterbankHeader::tstart},     {"tsamp", &FilterbankHeader::tsamp},
      {"period", &FilterbankHeader::period},     {"fch1", &FilterbankHeader::fch1},
      {"foff", &FilterbankHeader::foff},         {"refdm", &FilterbankHeader::refdm}};
  bool in_table = false;
  for (;;) {
    const std::string key = read_string(f);
    if (key == "HEADER_END") break;
    if (key == "rawdatafile") {
      h.rawdatafile = read_string(f);
    } else if (key == "source_name") {
      h.source_name = read_string(f);
    } else if (key == "FREQUENCY_START") {
      in_table = true;
      h.uses_frequency_table = true;
      h.frequency_table.clear();
    } else if (key == "FREQUENCY_END") {
      in_table = false;
      h.nchans = static_cast<int>(h.frequency_table.size());
    } else if (key == "fchannel") {
      if (!in_table) throw std::runtime_error("fchannel outside FREQUENCY block");
      h.frequency_table.push_back(read_value<double>(f, "fchannel"));
      h.fch1 = 0.0;
      h.foff = 0.0;
    } else if (key == "nsamples") {
      (void)read_value<std::int32_t>(f, "nsamples");  // recomputed from the file size
    } else {
      bool known = false;
      for (const IntKey& k : ints)
        if (key == k.key) {
          h.*(k.field) = read_value<std::int32_t>(f, k.key);
          known = true;
          break;
        }
      for (const DoubleKey& k : doubles)
        if (!known && key == k.key) {
          h.*(k.field) = read_value<double>(f, k.key);
          known = true;
          break;
        }
      if (!known) throw std::runtime_error("Unknown filterbank header field: " + key);
    }
  }
  const long end = std::ftell(f);
  if (end < 0) throw std::runtime_error("Failed to determine filterbank header size");
  h.header_size = static_cast<std::size_t>(end);
  if (h.nchans <= 0) throw std::runtime_error("Invalid filterbank nchans");
  if (h.nifs <= 0) throw std::runtime_error("Invalid filterbank nifs");
  if (h.nbits != 8 && h.nbits != 16 && h.nbits != 32)
    throw std::runtime_error("Unsupported filterbank nbits");
  if (std::fseek(f, 0, SEEK_END) != 0) throw std::runtime_error("Failed to size filterbank file");
  const long size = std::ftell(f);
  if (size < 0) throw std::runtime_error("Failed to size filterbank file");
  if (static_cast<std::size_t>(size) < h.header_size)
    throw std::runtime_error("Filterbank header is larger than file");
  const std::size_t data_bytes = static_cast<std::size_t>(size) - h.header_size;
  const std::size_t row =
      static_cast<std::size_t>(h.nifs) * static_cast<std::size_t>(h.nchans) * (h.nbits / 8);
  if (data_bytes % row != 0)
    throw std::runtime_error("Filterbank data size is not a whole number of rows");
  h.nsamples = data_bytes / row;
  if (h.uses_frequency_table) {
    if (h.frequency_table.size() != static_cast<std::size_t>(h.nchans))
      throw std::runtime_error("Frequency table length does not match nchans");
  } else {
    h.frequency_table.resize(static_cast<std::size_t>(h.nchans));
    for (int c = 0; c < h.nchans; ++c)
      h.frequency_table[static_cast<std::size_t>(c)] = h.fch1 + static_cast<double>(c) * h.foff;
  }
  const std::vector<double>& ft = h.frequency_table;
  const bool descending = h.uses_frequency_table ? (ft.size() >= 2 && ft.front() > ft.back())
                                                 : h.foff < 0.0;
  h.reversed = order == ChannelOrderPolicy::FrequencyAscending && descending;
  if (h.reversed) {
    std::reverse(h.frequency_table.begin(), h.frequency_table.end());
    h.fch1 = h.frequency_table.front();
    if (!h.uses_frequency_table) h.foff = std::abs(h.foff);
  }
  return h;
}

template <typename T>
void reverse_channel_rows(const T* src, T* dst, std::size_t rows, std::size_t nchans,
                          bool parallel) {
#pragma omp parallel for if (parallel) schedule(static)
  for (std::int64_t r = 0; r < static_cast<std::int64_t>(rows); ++r) {
    const T* a = src + static_cast<std::size_t>(r) * nchans;
    T* b = dst + static_cast<std::size_t>(r) * nchans;
    for (std::size_t c = 0; c < nchans; ++c) b[c] = a[nchans - 1 - c];
  }
}

}  // namespace

FilterbankHeader read_filterbank_header(const std::string& path, ChannelOrderPolicy order) {
  File file(path, "rb");
  return parse_header(file.f, order);
}

FilterbankData read_filterbank(const std::string& path, const FilterbankReadOptions& o) {
  File file(path, "rb");
  FilterbankData d;
  d.header = parse_header(file.f, o.channel_order);
  const FilterbankHeader& h = d.header;
  const std::size_t esize = static_cast<std::size_t>(h.nbits / 8);
  const std::size_t row_elems = static_cast<std::size_t>(h.nifs) * static_cast<std::size_t>(h.nchans);
  const std::size_t nchans = static_cast<std::size_t>(h.nchans);
  if (h.nsamples != 0 && row_elems * esize > std::numeric_limits<std::size_t>::max() / h.nsamples)
    throw std::overflow_error("Filterbank size overflow: total elements");
  const std::size_t total = h.nsamples * row_elems * esize;
  d.bytes.resize(total);
  if (std::fseek(file.f, static_cast<long>(h.header_size), SEEK_SET) != 0)
    throw std::runtime_error("Failed to read filterbank samples");
  if (!h.reversed) {
    if (total != 0 && std::fread(d.bytes.data(), 1, total, file.f) != total)
      throw std::runtime_error("Failed to read filterbank samples");
    return d;
  }
  // reversed: read io_buffer_bytes of whole rows at a time and write each
  // (row, if) block's channels back to front
  const std::size_t row_bytes = row_elems * esize;
  const std::size_t rows_per_chunk = std::max<std::size_t>(1, o.io_buffer_bytes / row_bytes);
  std::vector<std::uint8_t> buf(rows_per_chunk * row_bytes);
  const bool parallel = o.reverse_backend == ReverseBackend::CpuOpenmp ||
                        (o.reverse_backend == ReverseBackend::Auto &&
                         h.nsamples >= o.openmp_min_rows);
  for (std::size_t r0 = 0; r0 < h.nsamples; r0 += rows_per_chunk) {
    const std::size_t rows = std::min(rows_per_chunk, h.nsamples - r0);
    if (std::fread(buf.data(), 1, rows * row_bytes, file.f) != rows * row_bytes)
      throw std::runtime_error("Failed to read filterbank samples");
    const std::size_t blocks = rows * static_cast<std::size_t>(h.nifs);
    std::uint8_t* dst = d.bytes.data() + r0 * row_bytes;
    switch (esize) {
      case 1:
        reverse_channel_rows(buf.data(), dst, blocks, nchans, parallel);
        break;
      case 2:
        reverse_channel_rows(reinterpret_cast<const std::uint16_t*>(buf.data()),
                             reinterpret_cast<std::uint16_t*>(dst), blocks, nchans, parallel);
        break;
      default:
        reverse_channel_rows(reinterpret_cast<const std::uint32_t*>(buf.data()),
                             reinterpret_cast<std::uint32_t*>(dst), blocks, nchans, parallel);
        break;
    }
  }
  return d;
}

void write_filterbank_u8(const std::string& path, const FilterbankHeader& h,
                         const std::uint8_t* samples, std::size_t nsamples) {
  File file(path, "wb");
  std::FILE* f = file.f;
  write_string(f, "HEADER_START");
  if (!h.source_name.empty()) {
    write_string(f, "source_name");
    write_string(f, h.source_name);
  }
  write_field<std::int32_t>(f, "telescope_id", h.telescope_id);
  write_field<std::int32_t>(f, "machine_id", h.machine_id);
  write_field<std::int32_t>(f, "data_type", h.data_type);
  write_field<double>(f, "tstart", h.tstart);
  write_field<double>(f, "tsamp", h.tsamp);
  write_field<double>(f, "fch1", h.fch1);
  write_field<double>(f, "foff", h.foff);
  write_field<std::int32_t>(f, "nchans", h.nchans);
  write_field<std::int32_t>(f, "nbits", 8);
  write_field<std::int32_t>(f, "nifs", 1);
  write_string(f, "HEADER_END");
  const std::size_t bytes = nsamples * static_cast<std::size_t>(h.nchans);
  if (std::fwrite(samples, 1, bytes, f) != bytes)
    throw std::runtime_error("filterbank write failed: " + path);
}

FilterbankStreamStats stream_filterbank_u8(const std::string& path, const FilterbankHeader& h,
                                           std::uint8_t* device, gaffa_hip_stream compute,
                                           std::size_t chunk_samples,
                                           const std::function<void(std::size_t)>& on_chunk) {
  if (h.nbits != 8 || h.nifs != 1)
    throw std::invalid_argument("streamed upload supports 8-bit, nifs == 1 filterbanks");
  if (chunk_samples == 0) throw std::invalid_argument("chunk_samples must be positive");
  const double t_start = now_ms();
  const std::size_t nchans = static_cast<std::size_t>(h.nchans);
  const std::size_t chunk_bytes = chunk_samples * nchans;
  File file(path, "rb");
  std::fseek(file.f, static_cast<long>(h.header_size), SEEK_SET);
  // two pinned host buffers, two device staging buffers, one copy stream
  GpuPinnedBuffer pinned[2] = {GpuPinnedBuffer(chunk_bytes), GpuPinnedBuffer(chunk_bytes)};
  GpuDeviceBuffer<std::uint8_t> staging[2] = {GpuDeviceBuffer<std::uint8_t>(chunk_bytes),
                                              GpuDeviceBuffer<std::uint8_t>(chunk_bytes)};
  GpuStream copy = GpuStream::create();
  GpuEvent uploaded[2], consumed[2];
  FilterbankStreamStats st;
  for (std::size_t t0 = 0, k = 0; t0 < h.nsamples; t0 += chunk_samples, ++k) {
    const std::size_t rows = std::min(chunk_samples, h.nsamples - t0);
    const std::size_t b = k & 1u;
    // the pinned buffer's previous upload must be done before it is refilled
    if (k >= 2) uploaded[b].synchronize();
    const double r0 = now_ms();
    if (std::fread(pinned[b].data(), 1, rows * nchans, file.f) != rows * nchans)
      throw std::runtime_error("filterbank file ended early: " + path);
    st.read_ms += now_ms() - r0;
    // the staging buffer's previous contents must have been consumed
    if (k >= 2) consumed[b].wait(copy.get());
    hip_ok(gaffa_hip_memcpy(staging[b].data(), pinned[b].data(), rows * nchans, copy.get()),
           "hipMemcpyAsync");
    uploaded[b].record(copy.get());
    uploaded[b].wait(compute);
    hip_ok(gaffa_hip_copy_rows_u8(staging[b].data(), device + t0 * nchans, rows, nchans,
                                  h.reversed ? 1 : 0, compute),
           "copy_rows_u8");
    consumed[b].record(compute);
    st.bytes += rows * nchans;
    ++st.chunks;
    if (on_chunk) on_chunk(t0 + rows);
  }
  hip_ok(gaffa_hip_stream_synchronize(compute), "hipStreamSynchronize");
  st.wall_ms = now_ms() - t_start;
  return st;
}

}  // namespace gaffa
