// SIGPROC filterbank header parsing and streamed device upload.
#include "gaffa/filterbank.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
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
T read_value(std::FILE* f) {
  T v{};
  if (std::fread(&v, sizeof(T), 1, f) != 1)
    throw std::runtime_error("Non-standard filterbank file: truncated header");
  return v;
}

std::string read_string(std::FILE* f) {
  const std::int32_t n = read_value<std::int32_t>(f);
  if (n < 0 || n > 4096) throw std::runtime_error("Non-standard filterbank file: bad header string length");
  std::string s(static_cast<std::size_t>(n), '\0');
  if (n > 0 && std::fread(s.data(), 1, s.size(), f) != s.size())
    throw std::runtime_error("Non-standard filterbank file: truncated header");
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

FilterbankHeader read_filterbank_header(const std::string& path) {
  File file(path, "rb");
  FilterbankHeader h;
  if (read_string(file.f) != "HEADER_START")
    throw std::runtime_error("Non-standard filterbank file: missing HEADER_START");
  static const std::set<std::string> ints = {
      "telescope_id", "machine_id", "data_type", "barycentric", "pulsarcentric", "ibeam",
      "nbeams",       "npuls",      "nbins",     "nbits",       "nifs",          "nchans",
      "nsamples"};
  std::vector<double> table;
  for (;;) {
    const std::string key = read_string(file.f);
    if (key == "HEADER_END") break;
    if (key == "rawdatafile") {
      h.rawdatafile = read_string(file.f);
    } else if (key == "source_name") {
      h.source_name = read_string(file.f);
    } else if (ints.count(key) != 0) {
      const int v = read_value<std::int32_t>(file.f);
      if (key == "telescope_id") h.telescope_id = v;
      else if (key == "machine_id") h.machine_id = v;
      else if (key == "data_type") h.data_type = v;
      else if (key == "barycentric") h.barycentric = v;
      else if (key == "pulsarcentric") h.pulsarcentric = v;
      else if (key == "ibeam") h.ibeam = v;
      else if (key == "nbeams") h.nbeams = v;
      else if (key == "npuls") h.npuls = v;
      else if (key == "nbins") h.nbins = v;
      else if (key == "nbits") h.nbits = v;
      else if (key == "nifs") h.nifs = v;
      else if (key == "nchans") h.nchans = v;
      // nsamples: recomputed from the file size
    } else if (key == "fchannel") {
      table.push_back(read_value<double>(file.f));
    } else if (key == "FREQUENCY_START" || key == "FREQUENCY_END") {
      h.uses_frequency_table = true;
    } else {
      double* dst = key == "az_start"  ? &h.az_start
                    : key == "za_start" ? &h.za_start
                    : key == "src_raj"  ? &h.src_raj
                    : key == "src_dej"  ? &h.src_dej
                    : key == "tstart"   ? &h.tstart
                    : key == "tsamp"    ? &h.tsamp
                    : key == "fch1"     ? &h.fch1
                    : key == "foff"     ? &h.foff
                    : key == "refdm"    ? &h.refdm
                    : key == "period"   ? &h.period
                                        : nullptr;
      if (dst == nullptr) throw std::runtime_error("Unknown filterbank header field: " + key);
      *dst = read_value<double>(file.f);
    }
  }
  h.header_size = static_cast<std::size_t>(std::ftell(file.f));
  if (h.uses_frequency_table) {
    h.nchans = static_cast<int>(table.size());
  } else {
    for (int c = 0; c < h.nchans; ++c) table.push_back(h.fch1 + c * h.foff);
  }
  if (h.nchans <= 0) throw std::runtime_error("Invalid filterbank nchans");
  if (h.nbits != 8 && h.nbits != 16 && h.nbits != 32)
    throw std::runtime_error("Unsupported filterbank nbits");
  if (h.nifs <= 0) throw std::runtime_error("Invalid filterbank nifs");
  std::fseek(file.f, 0, SEEK_END);
  const std::size_t size = static_cast<std::size_t>(std::ftell(file.f));
  const std::size_t row = static_cast<std::size_t>(h.nifs) * h.nchans * (h.nbits / 8);
  h.nsamples = (size - h.header_size) / row;
  const bool descending = table.size() >= 2 && table.front() > table.back();
  h.reversed = h.uses_frequency_table ? descending : h.foff < 0.0;
  if (h.reversed) {
    std::reverse(table.begin(), table.end());
    h.fch1 = table.front();
    if (!h.uses_frequency_table) h.foff = std::abs(h.foff);
  }
  h.frequency_table = std::move(table);
  return h;
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
