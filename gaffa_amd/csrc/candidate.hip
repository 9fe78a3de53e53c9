// Candidate selection of a DM search on gfx950: gaffa's two-stage
// single-linkage clustering of peaks (candidate.cpp:175-212) over the
// compact peak records the FFA detection kernel wrote, so the 10^7-10^8
// peaks of a large search never have to be clustered on the host.
//
//   ranks     each peak's dense trial-period rank (host table per evaluated
//             (task, row)) and a flag per rank present in the search; the
//             host turns the flags into frequency clusters (gaps wider than
//             the radius split), which need the present ranks in frequency
//             order and are a few 10^5 entries
//   sort      a 64-bit key per peak, (frequency cluster, DM and width,
//             descending rank = ascending frequency), radix-sorted with its
//             peak index: the order select_candidates_cpu reaches with its
//             frequency sort and per-cluster (DM, width) sorts
//   split     a DM cluster starts at a new frequency cluster, a new width
//             (widths kept apart), or a DM gap wider than the radius
//   reduce    per DM cluster: frequency-rank and DM range, top S/N, count
//   ties      every peak whose S/N equals its cluster's top, for the host to
//             pick the representative with gaffa's full peak order
// The clusters, their ranges and their S/N maxima do not depend on the order
// of peaks with equal keys, so the result equals the CPU selection exactly.
#include "common.hpp"

#include <cmath>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_reduce_by_key.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

namespace gaffa_hip {
namespace {

constexpr uint32_t kThreads = 256;

__global__ void __launch_bounds__(kThreads)
candidate_rank_kernel(const gaffa_hip_ffa_peak* __restrict__ peaks, uint64_t n,
                      const uint32_t* __restrict__ period_rank,
                      const uint64_t* __restrict__ task_first_row, uint32_t ntasks,
                      uint32_t nranks, uint32_t ndm, uint32_t nwidths,
                      uint32_t* __restrict__ peak_rank, uint8_t* __restrict__ present,
                      uint32_t* __restrict__ status) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const gaffa_hip_ffa_peak p = peaks[i];
  uint32_t bad = 0;
  uint32_t r = 0;
  if (p.task < ntasks) {
    const uint64_t first = task_first_row[p.task];
    if (first + p.shift < task_first_row[p.task + 1]) r = period_rank[first + p.shift];
    else bad |= GAFFA_HIP_CANDIDATE_BAD_PEAK;
  } else {
    bad |= GAFFA_HIP_CANDIDATE_BAD_PEAK;
  }
  if (r >= nranks || p.series >= ndm || (p.phase_width >> 16) >= nwidths)
    bad |= GAFFA_HIP_CANDIDATE_BAD_PEAK;
  if (!isfinite(p.snr)) bad |= GAFFA_HIP_CANDIDATE_NONFINITE;
  if (bad != 0) {
    atomicOr(status, bad);
    r = 0;
  }
  peak_rank[i] = r;
  present[r] = 1;  // racing stores of the same value
}

struct KeyLayout {
  uint32_t rank_bits;   // low field: nranks - 1 - rank
  uint32_t major_bits;  // middle field: dm * nwidths + width (across widths)
                        //               or width * ndm + dm
  uint32_t nranks, ndm, nwidths;
  uint32_t across;
};

__device__ __forceinline__ uint64_t field(uint64_t key, uint32_t shift, uint32_t bits) {
  return (key >> shift) & ((uint64_t{1} << bits) - 1);
}

__global__ void __launch_bounds__(kThreads)
candidate_key_kernel(const gaffa_hip_ffa_peak* __restrict__ peaks, uint64_t n,
                     const uint32_t* __restrict__ peak_rank,
                     const uint32_t* __restrict__ cluster_of_rdesc, KeyLayout L,
                     uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (i >= n) return;
  const gaffa_hip_ffa_peak p = peaks[i];
  const uint32_t rdesc = L.nranks - 1 - peak_rank[i];
  const uint64_t w = p.phase_width >> 16;
  const uint64_t dm = p.series;
  const uint64_t major = L.across ? dm * L.nwidths + w : w * L.ndm + dm;
  keys[i] = (static_cast<uint64_t>(cluster_of_rdesc[rdesc]) << (L.major_bits + L.rank_bits)) |
            (major << L.rank_bits) | rdesc;
  vals[i] = static_cast<uint32_t>(i);
}

__global__ void __launch_bounds__(kThreads)
candidate_split_kernel(const uint64_t* __restrict__ keys, uint64_t n, KeyLayout L,
                       uint32_t dm_radius, uint32_t* __restrict__ head) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= n) return;
  if (k == 0) {
    head[k] = 1;
    return;
  }
  const uint64_t a = keys[k - 1], b = keys[k];
  const uint32_t cshift = L.major_bits + L.rank_bits;
  const uint64_t ma = field(a, L.rank_bits, L.major_bits);
  const uint64_t mb = field(b, L.rank_bits, L.major_bits);
  const uint64_t dma = L.across ? ma / L.nwidths : ma % L.ndm;
  const uint64_t dmb = L.across ? mb / L.nwidths : mb % L.ndm;
  const bool same_width = L.across || ma / L.ndm == mb / L.ndm;
  const bool same_dm = dmb <= dma || dmb - dma <= dm_radius;
  head[k] = ((a >> cshift) != (b >> cshift) || !same_width || !same_dm) ? 1u : 0u;
}

// One peak's contribution to its DM cluster's summary.
using Summary = gaffa_hip_candidate_segment;

struct SummaryOf {
  const uint64_t* keys;
  const uint32_t* vals;
  const gaffa_hip_ffa_peak* peaks;
  KeyLayout L;
  __device__ Summary operator()(uint64_t k) const {
    const uint64_t key = keys[k];
    const uint32_t rdesc = static_cast<uint32_t>(field(key, 0, L.rank_bits));
    const uint64_t major = field(key, L.rank_bits, L.major_bits);
    const uint32_t dm = static_cast<uint32_t>(L.across ? major / L.nwidths : major % L.ndm);
    return Summary{rdesc, rdesc, dm, dm, peaks[vals[k]].snr, 1u};
  }
};

struct Merge {
  __device__ Summary operator()(const Summary& a, const Summary& b) const {
    return Summary{min(a.rdesc_min, b.rdesc_min), max(a.rdesc_max, b.rdesc_max),
                   min(a.dm_min, b.dm_min), max(a.dm_max, b.dm_max),
                   a.snr > b.snr ? a.snr : b.snr, a.count + b.count};
  }
};

__global__ void __launch_bounds__(kThreads)
candidate_tie_kernel(const uint32_t* __restrict__ vals, const uint32_t* __restrict__ segment,
                     uint64_t n, const gaffa_hip_ffa_peak* __restrict__ peaks,
                     const gaffa_hip_candidate_segment* __restrict__ segments,
                     gaffa_hip_candidate_tie* __restrict__ ties, uint32_t* __restrict__ nties) {
  const uint64_t k = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = segment[k] - 1;  // inclusive scan of the head flags
  const uint32_t i = vals[k];
  if (peaks[i].snr == segments[s].snr) {
    const uint32_t slot = atomicAdd(nties, 1u);
    ties[slot] = gaffa_hip_candidate_tie{s, i};
  }
}

__global__ void fill_bytes_kernel(uint8_t* p, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kThreads)
    p[i] = 0;
}

uint32_t bits_for(uint64_t count) {  // bits to hold 0 .. count - 1 (at least 1)
  uint32_t b = 1;
  while (b < 64 && (uint64_t{1} << b) < count) ++b;
  return b;
}

// Workspace carve-up for n peaks: sorted and unsorted keys and values, head
// flags, cluster indices and their unique copies, then the rocPRIM temporary
// storage.
struct Carve {
  uint64_t *keys_in, *keys_out;
  uint32_t *vals_in, *vals_out, *head, *segment;
  uint32_t* unique;
  void* temp;
  size_t temp_bytes;
  uint64_t total;
};

constexpr uint64_t align256(uint64_t v) { return (v + 255) / 256 * 256; }

size_t temp_bytes_for(uint64_t n) {
  size_t sort = 0, scan = 0, reduce = 0;
  hip_check(rocprim::radix_sort_pairs(nullptr, sort, static_cast<uint64_t*>(nullptr),
                                      static_cast<uint64_t*>(nullptr),
                                      static_cast<uint32_t*>(nullptr),
                                      static_cast<uint32_t*>(nullptr), n, 0, 64),
            "radix_sort_pairs size");
  hip_check(rocprim::inclusive_scan(nullptr, scan, static_cast<uint32_t*>(nullptr),
                                    static_cast<uint32_t*>(nullptr), n,
                                    rocprim::plus<uint32_t>()),
            "inclusive_scan size");
  const SummaryOf of{nullptr, nullptr, nullptr, KeyLayout{}};
  auto values = rocprim::make_transform_iterator(rocprim::make_counting_iterator<uint64_t>(0), of);
  hip_check(rocprim::reduce_by_key(nullptr, reduce, static_cast<uint32_t*>(nullptr), values, n,
                                   static_cast<uint32_t*>(nullptr),
                                   static_cast<Summary*>(nullptr),
                                   static_cast<uint32_t*>(nullptr), Merge()),
            "reduce_by_key size");
  return std::max(sort, std::max(scan, reduce));
}

Carve carve(void* base, uint64_t n) {
  Carve c{};
  auto* p = static_cast<uint8_t*>(base);
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    uint8_t* q = p == nullptr ? nullptr : p + off;
    off += align256(bytes);
    return q;
  };
  c.keys_in = reinterpret_cast<uint64_t*>(take(8 * n));
  c.keys_out = reinterpret_cast<uint64_t*>(take(8 * n));
  c.vals_in = reinterpret_cast<uint32_t*>(take(4 * n));
  c.vals_out = reinterpret_cast<uint32_t*>(take(4 * n));
  c.head = reinterpret_cast<uint32_t*>(take(4 * n));
  c.segment = reinterpret_cast<uint32_t*>(take(4 * n));
  c.unique = reinterpret_cast<uint32_t*>(take(4 * n));
  c.temp_bytes = temp_bytes_for(n);
  c.temp = take(c.temp_bytes);
  c.total = off;
  return c;
}

dim3 grid_for(uint64_t n) { return dim3(static_cast<uint32_t>(ceil_div(n, kThreads))); }

}  // namespace
}  // namespace gaffa_hip

using namespace gaffa_hip;

extern "C" {

int gaffa_hip_candidate_workspace_bytes(uint64_t npeaks, uint64_t* bytes) {
  return guarded([&] {
    check(bytes != nullptr, "candidate workspace size pointer must not be null");
    check(npeaks > 0 && npeaks < (uint64_t{1} << 32), "candidate selection takes 1..2^32-1 peaks");
    *bytes = carve(nullptr, npeaks).total;
  });
}

int gaffa_hip_candidate_ranks(const gaffa_hip_candidate_args* a, gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr && a->peaks != nullptr && a->period_rank != nullptr &&
              a->task_first_row != nullptr && a->peak_rank != nullptr &&
              a->present != nullptr && a->status != nullptr,
          "candidate selection buffers must not be null");
    check(a->npeaks > 0 && a->npeaks < (uint64_t{1} << 32),
          "candidate selection takes 1..2^32-1 peaks");
    check(a->ntasks > 0 && a->nranks > 0 && a->ndm > 0 && a->nwidths > 0,
          "candidate selection needs tasks, ranks, DM trials and widths");
    const hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(fill_bytes_kernel, dim3(256), dim3(kThreads), 0, st, a->present,
                       static_cast<uint64_t>(a->nranks));
    launch_check("fill_bytes_kernel");
    hipLaunchKernelGGL(fill_bytes_kernel, dim3(1), dim3(kThreads), 0, st,
                       reinterpret_cast<uint8_t*>(a->status), uint64_t{4});
    launch_check("fill_bytes_kernel");
    hipLaunchKernelGGL(candidate_rank_kernel, grid_for(a->npeaks), dim3(kThreads), 0, st,
                       a->peaks, a->npeaks, a->period_rank, a->task_first_row, a->ntasks,
                       a->nranks, a->ndm, a->nwidths, a->peak_rank, a->present, a->status);
    launch_check("candidate_rank_kernel");
  });
}

int gaffa_hip_candidate_clusters(const gaffa_hip_candidate_args* a, gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr && a->peaks != nullptr && a->peak_rank != nullptr &&
              a->cluster_of_rdesc != nullptr && a->workspace != nullptr &&
              a->segments != nullptr && a->nsegments != nullptr && a->ties != nullptr &&
              a->nties != nullptr,
          "candidate selection buffers must not be null");
    const uint64_t n = a->npeaks;
    check(n > 0 && n < (uint64_t{1} << 32), "candidate selection takes 1..2^32-1 peaks");
    check(a->nranks > 0 && a->ndm > 0 && a->nwidths > 0 && a->nclusters > 0,
          "candidate selection needs ranks, DM trials, widths and clusters");
    KeyLayout L{};
    L.rank_bits = bits_for(a->nranks);
    L.major_bits = bits_for(static_cast<uint64_t>(a->ndm) * a->nwidths);
    L.nranks = a->nranks;
    L.ndm = a->ndm;
    L.nwidths = a->nwidths;
    L.across = a->cluster_across_widths != 0;
    const uint32_t total_bits = L.rank_bits + L.major_bits + bits_for(a->nclusters);
    check_overflow(total_bits <= 64, "candidate sort key exceeds 64 bits");
    Carve c = carve(a->workspace, n);
    check(c.total <= a->workspace_bytes, "candidate workspace too small");
    const hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(candidate_key_kernel, grid_for(n), dim3(kThreads), 0, st, a->peaks, n,
                       a->peak_rank, a->cluster_of_rdesc, L, c.keys_in, c.vals_in);
    launch_check("candidate_key_kernel");
    size_t bytes = c.temp_bytes;
    hip_check(rocprim::radix_sort_pairs(c.temp, bytes, c.keys_in, c.keys_out, c.vals_in,
                                        c.vals_out, n, 0, total_bits, st),
              "radix_sort_pairs");
    hipLaunchKernelGGL(candidate_split_kernel, grid_for(n), dim3(kThreads), 0, st, c.keys_out,
                       n, L, a->dm_cluster_radius, c.head);
    launch_check("candidate_split_kernel");
    bytes = c.temp_bytes;
    hip_check(rocprim::inclusive_scan(c.temp, bytes, c.head, c.segment, n,
                                      rocprim::plus<uint32_t>(), st),
              "inclusive_scan");
    const SummaryOf of{c.keys_out, c.vals_out, a->peaks, L};
    auto values = rocprim::make_transform_iterator(rocprim::make_counting_iterator<uint64_t>(0),
                                                   of);
    bytes = c.temp_bytes;
    hip_check(rocprim::reduce_by_key(c.temp, bytes, c.segment, values, n, c.unique, a->segments,
                                     a->nsegments, Merge(), rocprim::equal_to<uint32_t>(), st),
              "reduce_by_key");
    hipLaunchKernelGGL(fill_bytes_kernel, dim3(1), dim3(kThreads), 0, st,
                       reinterpret_cast<uint8_t*>(a->nties), uint64_t{4});
    launch_check("fill_bytes_kernel");
    hipLaunchKernelGGL(candidate_tie_kernel, grid_for(n), dim3(kThreads), 0, st, c.vals_out,
                       c.segment, n, a->peaks, a->segments, a->ties, a->nties);
    launch_check("candidate_tie_kernel");
  });
}

}  // extern "C"
