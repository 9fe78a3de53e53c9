// Incoherent dedispersion of uint8 filterbanks on gfx950.
//
// Brute force stages delay-shifted input through LDS: a workgroup copies a
// time slab of a channel block from HBM (row segments of contiguous channel
// bytes, coalesced), transposes it to channel-major bytes in LDS, then every
// thread reads 4 consecutive time samples of a channel with one aligned dword
// pair + v_alignbyte and accumulates them as packed 16-bit pairs (<= 32
// channels per flush, so no 16-bit lane can overflow).  Subband dedispersion
// reads a channel-major copy of the filterbank (transposed once, shared by
// every DM segment) directly from global memory, 8 samples per lane, and
// keeps each nominal group's subband sums in a sliding LDS window.
// Integer sums are order-independent, so results equal gaffa's CPU sums bit
// for bit (dedispersion.cpp:197-215, 291-329, 351-453).
#include "common.hpp"

#include <type_traits>

namespace gaffa_hip {
namespace {

constexpr int kThreads = 256;

// ---------------------------------------------------------- brute force ----
constexpr int kBfTime = 256;  // output samples per workgroup (64 lanes x 4)
constexpr int kBfDms = 16;    // DMs per workgroup (4 waves x 4)
constexpr int kBfChans = 32;  // channels per LDS slab

__device__ inline uint32_t read_bytes4(const uint32_t* slab32, uint32_t off) {
  // bytes [off, off + 4) of a byte array viewed as dwords
  const uint32_t a = off >> 2;
  const uint32_t lo = slab32[a];
  const uint32_t hi = slab32[a + 1];
  return __builtin_amdgcn_alignbyte(hi, lo, off & 3u);
}

__global__ void __launch_bounds__(kThreads)
dedisperse_multi_u8_kernel(const uint8_t* __restrict__ x, uint64_t nsamples,
                           uint64_t nchans, uint64_t chan_begin, uint32_t nsel,
                           const int32_t* __restrict__ delays, uint32_t ndm,
                           uint32_t* __restrict__ out, uint64_t out_nsamples,
                           uint64_t out_stride, uint32_t slab_len,
                           int32_t* __restrict__ error_flag) {
  extern __shared__ uint32_t dyn_lds[];
  uint8_t* slab = reinterpret_cast<uint8_t*>(dyn_lds);  // [kBfChans][slab_len]
  __shared__ int32_t sdel[kBfDms][kBfChans];
  __shared__ int32_t sminmax[2];

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = tid >> 6;
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kBfTime;
  const uint32_t dm0 = blockIdx.y * kBfDms;
  const uint32_t nd = min(static_cast<uint32_t>(kBfDms), ndm - dm0);

  uint32_t total[4][4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int j = 0; j < 4; ++j) total[k][j] = 0;

  for (uint32_t cb0 = 0; cb0 < nsel; cb0 += kBfChans) {
    const uint32_t ncb = min(static_cast<uint32_t>(kBfChans), nsel - cb0);
    if (tid == 0) {
      sminmax[0] = 0x7fffffff;
      sminmax[1] = 0;
    }
    __syncthreads();
    int32_t lmin = 0x7fffffff, lmax = 0;
    for (uint32_t e = tid; e < kBfDms * kBfChans; e += kThreads) {
      const uint32_t i = e / kBfChans, c = e % kBfChans;
      int32_t d = 0;
      if (i < nd && c < ncb) {
        d = delays[static_cast<uint64_t>(dm0 + i) * nsel + cb0 + c];
        lmin = min(lmin, d);
        lmax = max(lmax, d);
      }
      sdel[i][c] = d;
    }
    atomicMin(&sminmax[0], lmin);
    atomicMax(&sminmax[1], lmax);
    __syncthreads();
    const int32_t dmin = sminmax[0];
    const int32_t spread = sminmax[1] - dmin;
    // The host sized slab_len from its own table; a mismatch would overrun
    // LDS, so refuse the block (uniform branch) and report it.
    if (static_cast<uint32_t>(spread) + kBfTime + 8 > slab_len) {
      if (tid == 0) atomicExch(error_flag, 1);
      return;
    }
    // Stage the slab: rows [t0 + dmin, t0 + dmin + rows) x ncb channels.
    const uint32_t rows = kBfTime + spread + 8;
    for (uint32_t e = tid; e < rows * ncb; e += kThreads) {
      const uint32_t r = e / ncb, c = e % ncb;
      const uint64_t t = t0 + static_cast<uint64_t>(dmin) + r;
      uint8_t v = 0;
      if (t < nsamples) v = x[t * nchans + chan_begin + cb0 + c];
      slab[c * slab_len + r] = v;
    }
    __syncthreads();

    const uint32_t* slab32 = reinterpret_cast<const uint32_t*>(slab);
    uint32_t acc02[4] = {0, 0, 0, 0}, acc13[4] = {0, 0, 0, 0};
    for (uint32_t c = 0; c < ncb; ++c) {
      const uint32_t row = c * slab_len + 4u * lane;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const uint32_t i = wave + 4u * k;
        // padding DMs (i >= nd) hold delay 0, which may sit below dmin
        const int32_t rel = max(sdel[i][c] - dmin, 0);
        const uint32_t off = row + static_cast<uint32_t>(rel);
        const uint32_t w = read_bytes4(slab32, off);
        acc02[k] += w & 0x00FF00FFu;
        acc13[k] += (w >> 8) & 0x00FF00FFu;
      }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      total[k][0] += acc02[k] & 0xFFFFu;
      total[k][2] += acc02[k] >> 16;
      total[k][1] += acc13[k] & 0xFFFFu;
      total[k][3] += acc13[k] >> 16;
    }
    __syncthreads();
  }

#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = wave + 4u * k;
    if (i >= nd) continue;
    uint32_t* o = out + static_cast<uint64_t>(dm0 + i) * out_stride;
    const uint64_t t = t0 + 4u * lane;
    if (t + 3 < out_nsamples && (out_stride & 3u) == 0 &&
        (reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
      *reinterpret_cast<uint4*>(o + t) =
          make_uint4(total[k][0], total[k][1], total[k][2], total[k][3]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j)
        if (t + j < out_nsamples) o[t + j] = total[k][j];
    }
  }
}

// ---------------------------------------------------------- transpose ----
// out[c][t] = x[t][chan_begin + c] for t < nsamples, 0 for nsamples <= t <
// out_stride: one workgroup moves a 256-sample x 64-channel tile through LDS
// (16-byte loads along channels, byte writes into [channel][time] rows,
// 16-byte stores along time).
constexpr uint32_t kTrTime = 256;
constexpr uint32_t kTrChans = 64;
constexpr uint32_t kTrStride = kTrTime + 16;  // LDS row bytes (16-aligned)

__global__ void __launch_bounds__(kThreads)
transpose_u8_kernel(const uint8_t* __restrict__ x, uint64_t nsamples, uint64_t nchans,
                    uint64_t chan_begin, uint32_t nsel, uint8_t* __restrict__ out,
                    uint64_t out_stride, uint32_t vec_in, uint64_t t_begin) {
  __shared__ __align__(16) uint8_t tile[kTrChans * kTrStride];
  const uint32_t tid = threadIdx.x;
  const uint64_t t0 = t_begin + static_cast<uint64_t>(blockIdx.x) * kTrTime;
  const uint32_t c0 = blockIdx.y * kTrChans;
  const uint32_t nc = min(kTrChans, nsel - c0);
  // 256 rows x 4 chunks of 16 channels
#pragma unroll
  for (uint32_t r = 0; r < 4; ++r) {
    const uint32_t row = (tid >> 2) + 64u * r;
    const uint32_t chunk = tid & 3u;
    const uint64_t t = t0 + row;
    uint8_t v[16];
    if (vec_in && 16u * chunk + 16u <= nc) {
      uint4 q = make_uint4(0, 0, 0, 0);
      if (t < nsamples)
        q = *reinterpret_cast<const uint4*>(x + t * nchans + chan_begin + c0 + 16u * chunk);
      const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = static_cast<uint8_t>(w[u >> 2] >> (8 * (u & 3)));
    } else {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const uint32_t c = 16u * chunk + u;
        v[u] = (t < nsamples && c < nc) ? x[t * nchans + chan_begin + c0 + c] : 0;
      }
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) tile[(16u * chunk + u) * kTrStride + row] = v[u];
  }
  __syncthreads();
  // 64 channel rows x 16 chunks of 16 samples
#pragma unroll
  for (uint32_t r = 0; r < 4; ++r) {
    const uint32_t c = (tid >> 4) + 16u * r;
    const uint32_t part = tid & 15u;
    const uint64_t t = t0 + 16u * part;
    if (c < nc && t < out_stride) {
      const uint4 q = *reinterpret_cast<const uint4*>(tile + c * kTrStride + 16u * part);
      *reinterpret_cast<uint4*>(out + static_cast<uint64_t>(c0 + c) * out_stride + t) = q;
    }
  }
}

// ------------------------------------------------------------- subband ----
// One workgroup = (nominal DM group, time span).  The group's u16 subband
// intermediate is a sliding window in LDS: subband s keeps
// [t, t + chunk + h_s) where h_s is the largest residual delay of the
// group's DMs in that subband (rounded up to 8).  The workgroup walks its
// span chunk by chunk; each step
//   1. moves the last h_s samples of every window to its front (no longer
//      needed samples drop out),
//   2. stage 1: computes the `chunk` new intermediate samples of every
//      subband from the channel-major rows: lane w sums samples
//      [8w, 8w + 8) of each channel of the subband at the group's coarse
//      delay with one 12-byte load and four v_perm_b32 byte picks into
//      packed u16 pairs (<= 256 channels x 255 fits 16 bits),
//   3. stage 2: for each DM of the group, adds the subbands at their
//      residual delays (lane w: outputs [8w, 8w + 8), five LDS dwords per
//      subband, packed u16 pairs flushed to u32 before any half can
//      overflow) and stores the chunk of the DM's output row.
// The residual halo is therefore computed once per span, not once per
// chunk, and any residual span that fits the window works: no nominal
// group is narrowed.  A wave owns the windows of subbands wave,
// wave + waves, ...: its LDS operations run in order, so step 1 and step 2
// of one window need no barrier between them.  Workgroups are ordered span-
// major and group-minor over the XCDs, so the groups running together on one
// XCD read the same stretch of the channel rows.
// Threads per workgroup (a template argument, so loop strides stay
// constants): 1024 (16 waves) where the LDS windows leave room for one
// 512-thread workgroup per CU, else 512 (several per CU).
constexpr int kSbThreads = 1024;
constexpr int kSbThreadsSmall = 512;
constexpr uint32_t kSbPiece = 512;   // intermediate samples per wave pass (64 lanes x 8)
constexpr int kSbUnroll = 16;        // channels whose loads are in flight together
constexpr uint32_t kXcds = 8;        // MI355X: 8 XCDs, one L2 each
constexpr uint32_t kSbReadPad = 528; // stage-1 bytes a pass reads past its first sample

struct SubbandGroupDev {
  uint32_t coarse_row, first_dm, ndm, reserved;
};

__device__ inline uint32_t sb_sel_even(uint32_t sh) {
  // bytes sh and sh + 2 of {hi:lo} into the low bytes of two u16 halves
  return 0x0C000C00u | sh | ((sh + 2u) << 16);
}

// Stage-1 sums of N consecutive channels (rows `stride` bytes apart from
// rows0, delays kb[0 .. N)) at intermediate samples tb + 8 lane + [0, 8):
// every row load of the batch is issued before the adds.
template <int N>
__device__ inline void sb_channels(const uint8_t* rows0, uint64_t stride,
                                   const int32_t* __restrict__ kb, uint64_t tb, uint32_t lane,
                                   uint32_t& a0, uint32_t& a1, uint32_t& a2, uint32_t& a3) {
  uint32_t x0[N], x1[N], x2[N], sh[N];
  const uint8_t* rp = rows0;
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const uint64_t b = tb + static_cast<uint32_t>(kb[u]);
    const uint32_t* row = reinterpret_cast<const uint32_t*>(rp + (b & ~uint64_t{3}));
    sh[u] = static_cast<uint32_t>(b & 3u);
    x0[u] = row[2u * lane];
    x1[u] = row[2u * lane + 1u];
    x2[u] = row[2u * lane + 2u];
    rp += stride;
  }
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const uint32_t se = sb_sel_even(sh[u]);
    const uint32_t so = se + 0x00010001u;  // bytes sh + 1, sh + 3
    a0 += __builtin_amdgcn_perm(x1[u], x0[u], se);  // samples 0, 2
    a1 += __builtin_amdgcn_perm(x1[u], x0[u], so);  // samples 1, 3
    a2 += __builtin_amdgcn_perm(x2[u], x1[u], se);  // samples 4, 6
    a3 += __builtin_amdgcn_perm(x2[u], x1[u], so);  // samples 5, 7
  }
}

template <uint32_t kThreadsT>
__global__ void __launch_bounds__(kThreadsT)
dedisperse_subband_stream_kernel(const uint8_t* __restrict__ cm, uint64_t cm_stride,
                                 uint64_t cm_length, uint32_t nsel, uint32_t sbc, uint32_t nsub,
                                 const SubbandGroupDev* __restrict__ groups, uint32_t ngroups,
                                 uint32_t npn, uint32_t ndm, uint32_t n_coarse_rows,
                                 const int32_t* __restrict__ coarse,
                                 const int32_t* __restrict__ resid,
                                 uint32_t* __restrict__ out, uint64_t out_nsamples,
                                 uint64_t out_stride, uint32_t chunk, uint32_t span,
                                 uint32_t nspans, uint32_t max_group, uint32_t window_cap,
                                 uint32_t flush_every, int32_t* __restrict__ error_flag) {
  extern __shared__ __align__(16) uint32_t dyn_lds[];
  const uint32_t nsub4 = (nsub + 3u) & ~3u;
  // [max_group][nsub4] window addresses (u16 index) of each (DM, subband),
  // then [nsub4] window starts and [nsub4] halos, then the windows
  int32_t* saddr = reinterpret_cast<int32_t*>(dyn_lds);
  int32_t* soff = saddr + static_cast<size_t>(max_group) * nsub4;
  int32_t* shalo = soff + nsub4;
  uint16_t* win = reinterpret_cast<uint16_t*>(shalo + nsub4);
  uint32_t* win32 = reinterpret_cast<uint32_t*>(win);

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  constexpr uint32_t threads = kThreadsT;
  constexpr uint32_t nwaves = threads / 64u;
  const uint32_t nwork = ngroups * nspans;
  const uint32_t per_xcd = (nwork + kXcds - 1) / kXcds;
  const uint32_t w_id = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (w_id >= nwork) return;
  const uint32_t g = w_id % ngroups;
  const uint64_t t_begin = static_cast<uint64_t>(w_id / ngroups) * span;
  if (t_begin >= out_nsamples) return;
  const uint64_t t_end = min(t_begin + span, out_nsamples);
  uint32_t k, dm_first, nd;
  if (groups != nullptr) {
    k = groups[g].coarse_row;
    dm_first = groups[g].first_dm;
    nd = groups[g].ndm;
    // rows outside the tables: refuse the group (uniform)
    if (k >= n_coarse_rows || dm_first >= ndm || nd > ndm - dm_first) nd = 0;
  } else {
    k = g;
    dm_first = g * npn;
    nd = min(npn, ndm - dm_first);
  }

  __shared__ int32_t sbad;
  if (tid == 0) sbad = 0;
  __syncthreads();
  if (nd == 0 || nd > max_group) {
    if (tid == 0) atomicExch(error_flag, 1);
    return;
  }
  for (uint32_t e = tid; e < nd * nsub; e += threads) {
    const uint32_t i = e / nsub, s = e % nsub;
    const int32_t r = resid[static_cast<uint64_t>(dm_first + i) * nsub + s];
    if (r < 0 || r > (1 << 24)) sbad = 1;
    saddr[i * nsub4 + s] = r;
  }
  __syncthreads();
  for (uint32_t s = tid; s < nsub; s += threads) {
    int32_t h = 0;
    for (uint32_t i = 0; i < nd; ++i) h = max(h, saddr[i * nsub4 + s]);
    shalo[s] = (h + 7) & ~7;
  }
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (uint32_t s = 0; s < nsub; ++s) {
      soff[s] = static_cast<int32_t>(acc);
      acc += chunk + static_cast<uint32_t>(shalo[s]) + 8u;
    }
    if (acc > window_cap) sbad = 1;
  }
  __syncthreads();
  if (sbad) {
    if (tid == 0) atomicExch(error_flag, 1);
    return;
  }
  for (uint32_t e = tid; e < nd * nsub; e += threads) {
    const uint32_t i = e / nsub, s = e % nsub;
    saddr[i * nsub4 + s] += soff[s];
  }
  // every stage-1 read of this span stays inside the padded rows
  const uint32_t nchunks = static_cast<uint32_t>((t_end - t_begin + chunk - 1) / chunk);
  int32_t hmax = 0;
  for (uint32_t s = 0; s < nsub; ++s) hmax = max(hmax, shalo[s]);  // LDS broadcast reads
  const int64_t room = static_cast<int64_t>(cm_length) -
                       static_cast<int64_t>(t_begin + static_cast<uint64_t>(nchunks) * chunk) -
                       hmax - static_cast<int64_t>(kSbReadPad);
  const int32_t* kdel = coarse + static_cast<uint64_t>(k) * nsel;
  for (uint32_t c = tid; c < nsel; c += threads) {
    const int32_t d = kdel[c];
    if (d < 0 || static_cast<int64_t>(d) > room) sbad = 1;
  }
  __syncthreads();
  if (sbad) {
    if (tid == 0) atomicExch(error_flag, 1);
    return;
  }

  for (uint32_t j = 0; j < nchunks; ++j) {
    const uint64_t t_out = t_begin + static_cast<uint64_t>(j) * chunk;
    // Steps 1 and 2, per owned subband window.
    for (uint32_t s = wave; s < nsub; s += nwaves) {
      const uint32_t h = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(shalo[s]));
      const uint32_t base = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(soff[s]));
      if (j > 0) {
        // forward move in steps of `chunk`: a step's stores land on the
        // previous step's sources, which this wave has already read
        for (uint32_t x0 = 0; x0 < h; x0 += chunk) {
          const uint32_t x = x0 + 8u * lane;
          if (8u * lane < chunk && x < h) {
            const uint4 v = *reinterpret_cast<const uint4*>(win + base + chunk + x);
            *reinterpret_cast<uint4*>(win + base + x) = v;
          }
        }
      }
      asm volatile("" ::: "memory");  // the moves' LDS reads stay before stage-1 stores
      const uint32_t lo = j == 0 ? 0u : h;
      const uint32_t hi = h + chunk;
      const uint32_t cs = s * sbc;
      const uint32_t ce = min(cs + sbc, nsel);
      for (uint32_t p0 = lo; p0 < hi; p0 += kSbPiece) {
        const uint64_t tb = t_out + p0;  // intermediate time of the pass's first sample
        uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
        uint32_t c0 = cs;
        for (; c0 + kSbUnroll <= ce; c0 += kSbUnroll)
          sb_channels<kSbUnroll>(cm + static_cast<uint64_t>(c0) * cm_stride, cm_stride,
                                 kdel + c0, tb, lane, a0, a1, a2, a3);
        for (; c0 < ce; ++c0)
          sb_channels<1>(cm + static_cast<uint64_t>(c0) * cm_stride, cm_stride, kdel + c0, tb,
                         lane, a0, a1, a2, a3);
        if (p0 + 8u * lane < hi) {
          uint4 v;
          v.x = __builtin_amdgcn_perm(a1, a0, 0x05040100u);  // s0 | s1 << 16
          v.y = __builtin_amdgcn_perm(a1, a0, 0x07060302u);  // s2 | s3 << 16
          v.z = __builtin_amdgcn_perm(a3, a2, 0x05040100u);
          v.w = __builtin_amdgcn_perm(a3, a2, 0x07060302u);
          *reinterpret_cast<uint4*>(win + base + p0 + 8u * lane) = v;
        }
      }
    }
    __syncthreads();

    // Step 3: one wave pass per (DM, chunk outputs).  Lane w sums the output
    // pairs at 2w + k (chunk / 4), k < 4, so that for each k the lanes read
    // consecutive LDS dwords (a pair and the next dword for the odd-residual
    // shift: one ds_read2 per k, free of bank conflicts).  Lanes past
    // chunk / 8 in a short chunk re-read the last pair and store nothing.
    // The full chunk (512 samples) has its pair distance as a constant, so
    // the four reads share one address with immediate offsets.
    auto sum_dms = [&](auto slot_v) {
      const uint32_t slot = slot_v;  // dwords between a lane's pairs: chunk / 8
      const bool lane_live = lane < slot;
      const uint32_t wdw = min(lane, slot - 1u);
      for (uint32_t i = wave; i < nd; i += nwaves) {
        const int32_t* arow = saddr + i * nsub4;
        uint32_t tot[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        for (uint32_t s0 = 0; s0 < nsub; s0 += flush_every) {
          const uint32_t s1 = min(s0 + flush_every, nsub);
          uint32_t p0 = 0, p1 = 0, p2 = 0, p3 = 0;
#pragma unroll 4
          for (uint32_t s = s0; s < s1; ++s) {
            const uint32_t a = static_cast<uint32_t>(arow[s]);  // window u16 index
            const uint32_t* r32 = win32 + (a >> 1) + wdw;
            const uint32_t shb = (a & 1u) << 1;
            p0 += __builtin_amdgcn_alignbyte(r32[1], r32[0], shb);
            p1 += __builtin_amdgcn_alignbyte(r32[slot + 1u], r32[slot], shb);
            p2 += __builtin_amdgcn_alignbyte(r32[2u * slot + 1u], r32[2u * slot], shb);
            p3 += __builtin_amdgcn_alignbyte(r32[3u * slot + 1u], r32[3u * slot], shb);
          }
          tot[0] += p0 & 0xFFFFu;
          tot[1] += p0 >> 16;
          tot[2] += p1 & 0xFFFFu;
          tot[3] += p1 >> 16;
          tot[4] += p2 & 0xFFFFu;
          tot[5] += p2 >> 16;
          tot[6] += p3 & 0xFFFFu;
          tot[7] += p3 >> 16;
        }
        if (lane_live) {
          uint32_t* o = out + static_cast<uint64_t>(dm_first + i) * out_stride;
          const bool pairs =
              (out_stride & 1u) == 0 && (reinterpret_cast<uintptr_t>(out) & 7u) == 0;
#pragma unroll
          for (uint32_t k = 0; k < 4; ++k) {
            const uint64_t t = t_out + 2u * lane + 2u * slot * k;
            if (t + 2 <= t_end && pairs) {
              *reinterpret_cast<uint2*>(o + t) = make_uint2(tot[2 * k], tot[2 * k + 1]);
            } else {
              if (t < t_end) o[t] = tot[2 * k];
              if (t + 1 < t_end) o[t + 1] = tot[2 * k + 1];
            }
          }
        }
      }
    };
    if (chunk == kSbPiece) sum_dms(std::integral_constant<uint32_t, kSbPiece / 8>{});
    else sum_dms(chunk >> 3);
    __syncthreads();
  }
}

constexpr uint32_t kMaxLds = 160 * 1024;

int32_t* error_flag_for_device() {
  // one flag per device, created on first use and kept for the process
  static int32_t* flags[64] = {};
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= 64) throw status_error(GAFFA_HIP_RUNTIME, "device ordinal out of range");
  if (flags[dev] == nullptr) {
    int32_t* p = nullptr;
    hip_check(hipMalloc(&p, sizeof(int32_t)), "hipMalloc");
    hip_check(hipMemset(p, 0, sizeof(int32_t)), "hipMemset");
    flags[dev] = p;
  }
  return flags[dev];
}

}  // namespace

// LDS bytes of the subband kernel: the (DM, subband) address table, the
// window starts and halos, and the u16 windows (chunk + halo + 8 per subband).
static uint64_t subband_lds_bytes(uint64_t nsub, uint64_t max_group, uint64_t halo_sum,
                                  uint32_t chunk) {
  const uint64_t nsub4 = (nsub + 3) & ~3ull;
  return 4 * (max_group * nsub4 + 2 * nsub4) + 2 * (nsub * (chunk + 8ull) + halo_sum);
}

// Largest chunk (512 down to 8 samples) whose LDS fits one workgroup; 0 when
// none does.
static uint32_t subband_chunk(uint64_t nsub, uint64_t max_group, uint64_t halo_sum,
                              uint64_t& lds) {
  for (uint32_t chunk = kSbPiece; chunk >= 8; chunk /= 2) {
    lds = subband_lds_bytes(nsub, max_group, halo_sum, chunk);
    if (lds <= kMaxLds - 256) return chunk;
  }
  lds = 0;
  return 0;
}

static uint64_t subband_halo_sum(uint64_t nsub, int32_t max_residual, int64_t residual_window) {
  if (residual_window > 0) return static_cast<uint64_t>(residual_window);
  return nsub * ((static_cast<uint64_t>(std::max(max_residual, 0)) + 7) & ~7ull);
}

// Reads and clears the LDS-bound guard flag (a host/kernel table mismatch).
static void check_dedispersion_guard(hipStream_t stream) {
  int32_t* flag = error_flag_for_device();
  int32_t v = 0;
  hip_check(hipMemcpyAsync(&v, flag, sizeof(v), hipMemcpyDeviceToHost, stream), "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  if (v != 0) {
    hip_check(hipMemsetAsync(flag, 0, sizeof(int32_t), stream), "hipMemsetAsync");
    throw status_error(GAFFA_HIP_LOGIC, "dedispersion delay bound smaller than the delay table");
  }
}

}  // namespace gaffa_hip

using namespace gaffa_hip;

extern "C" {

int gaffa_hip_dedisperse_multi_u8(const gaffa_hip_dedisperse_multi_args* a,
                                  gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr, "dedispersion arguments must not be null");
    check(a->samples != nullptr && a->delays != nullptr && a->out != nullptr,
          "dedispersion device pointers must not be null");
    check(a->nsamples > 0, "dedispersion requires at least one sample");
    check(a->nchans > 0, "dedispersion requires at least one channel");
    check(a->chan_begin < a->chan_end, "channel range must be non-empty");
    check(a->chan_end <= a->nchans, "channel range exceeds sample channel count");
    check(a->ndm > 0, "dedispersion requires at least one DM");
    check(a->out_nsamples > 0 && a->out_nsamples <= a->nsamples,
          "dedispersion valid output range is empty");
    check(a->out_stride >= a->out_nsamples, "output stride must be >= out_nsamples");
    check(a->max_block_spread >= 0, "dedispersion requires non-negative delays");
    check(a->ndm <= 0xFFFFFFFFull && a->chan_end - a->chan_begin <= 0xFFFFFFFFull,
          "dedispersion table too large");
    const uint32_t slab_len =
        static_cast<uint32_t>((kBfTime + a->max_block_spread + 8 + 3) & ~3);
    const uint64_t lds = static_cast<uint64_t>(kBfChans) * slab_len;
    check_overflow(lds <= kMaxLds - 4096, "dedispersion delay spread too large for LDS");
    const uint64_t tiles = ceil_div(a->out_nsamples, kBfTime);
    check_overflow(tiles < 0x7FFFFFFFull, "dedispersion output too long");
    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(dedisperse_multi_u8_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds)),
              "hipFuncSetAttribute");
    const dim3 grid(static_cast<uint32_t>(tiles),
                    static_cast<uint32_t>(ceil_div(a->ndm, kBfDms)));
    hipLaunchKernelGGL(dedisperse_multi_u8_kernel, grid, dim3(kThreads),
                       static_cast<uint32_t>(lds), as_stream(stream), a->samples,
                       a->nsamples, a->nchans, a->chan_begin,
                       static_cast<uint32_t>(a->chan_end - a->chan_begin), a->delays,
                       static_cast<uint32_t>(a->ndm), a->out, a->out_nsamples,
                       a->out_stride, slab_len, error_flag_for_device());
    launch_check("dedisperse_multi_u8_kernel");
  });
}

int gaffa_hip_filterbank_transpose_u8(const gaffa_hip_transpose_args* a,
                                      gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr, "transpose arguments must not be null");
    check(a->samples != nullptr && a->out != nullptr, "transpose pointers must not be null");
    check(a->nsamples > 0 && a->nchans > 0, "transpose requires a non-empty filterbank");
    check(a->chan_begin < a->chan_end && a->chan_end <= a->nchans,
          "channel range exceeds sample channel count");
    check(a->out_stride >= a->nsamples && a->out_stride % 16 == 0,
          "channel-major stride must be >= nsamples and a multiple of 16");
    check((reinterpret_cast<uintptr_t>(a->out) & 15u) == 0,
          "channel-major output must be 16-byte aligned");
    const uint64_t nsel = a->chan_end - a->chan_begin;
    check_overflow(nsel <= 0xFFFFFFFFull, "transpose channel range too large");
    const uint64_t t_count = a->t_count == 0 ? a->out_stride - a->t_begin : a->t_count;
    check(a->t_begin % kTrTime == 0 && a->t_begin < a->out_stride &&
              t_count <= a->out_stride - a->t_begin &&
              (t_count % kTrTime == 0 || a->t_begin + t_count == a->out_stride),
          "transpose time range must be 256-aligned and inside the rows");
    const uint64_t tblocks = ceil_div(t_count, kTrTime);
    const uint64_t cblocks = ceil_div(nsel, kTrChans);
    check_overflow(tblocks < 0x7FFFFFFFull && cblocks < 65536, "transpose grid too large");
    const uint32_t vec_in = (a->nchans % 16 == 0 && a->chan_begin % 16 == 0 &&
                             (reinterpret_cast<uintptr_t>(a->samples) & 15u) == 0)
                                ? 1u
                                : 0u;
    hipLaunchKernelGGL(transpose_u8_kernel,
                       dim3(static_cast<uint32_t>(tblocks), static_cast<uint32_t>(cblocks)),
                       dim3(kThreads), 0, as_stream(stream), a->samples, a->nsamples, a->nchans,
                       a->chan_begin, static_cast<uint32_t>(nsel), a->out, a->out_stride,
                       vec_in, a->t_begin);
    launch_check("transpose_u8_kernel");
  });
}

int gaffa_hip_dedisperse_subband_u8(const gaffa_hip_dedisperse_subband_args* a,
                                    gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr, "dedispersion arguments must not be null");
    check(a->channels != nullptr && a->coarse_delays != nullptr &&
              a->residual_delays != nullptr && a->out != nullptr,
          "dedispersion device pointers must not be null");
    check(a->nsamples > 0, "dedispersion requires at least one sample");
    check(a->nsel > 0, "channel range must be non-empty");
    const uint64_t row_len = a->channel_length == 0 ? a->channel_stride : a->channel_length;
    check(row_len <= a->channel_stride || a->nsel == 1, "channel_length exceeds the row pitch");
    check(row_len >= a->nsamples && a->channel_stride % 4 == 0 &&
              (reinterpret_cast<uintptr_t>(a->channels) & 3u) == 0,
          "channel rows must be dword aligned and hold nsamples");
    check(a->ndm > 0, "dedispersion requires at least one DM");
    check(a->subband_channels > 0, "subband_channels must be positive");
    check(a->subband_channels <= 256,
          "GPU subband dedispersion supports at most 256 channels per subband");
    check(a->out_nsamples > 0 && a->out_nsamples <= a->nsamples,
          "dedispersion valid output range is empty");
    check(a->out_stride >= a->out_nsamples, "output stride must be >= out_nsamples");
    check(a->max_residual >= 0 && a->max_coarse >= 0,
          "dedispersion requires non-negative delays");
    const uint64_t nsel = a->nsel;
    check_overflow(nsel <= 0xFFFFFFFFull && a->ndm <= 0xFFFFFFFFull,
                   "dedispersion table too large");
    const uint64_t nsub = ceil_div(nsel, a->subband_channels);
    uint64_t ngroups = 0, max_group = 0, n_coarse = 0, npn = 0;
    if (a->groups != nullptr) {
      check(a->ngroups > 0 && a->max_group_ndm > 0 && a->n_coarse_rows > 0,
            "subband groups need ngroups, max_group_ndm and n_coarse_rows");
      ngroups = a->ngroups;
      max_group = a->max_group_ndm;
      n_coarse = a->n_coarse_rows;
    } else {
      check(a->ndm_per_nominal > 0, "ndm_per_nominal must be positive");
      npn = a->ndm_per_nominal;
      ngroups = ceil_div(a->ndm, npn);
      max_group = std::min<uint64_t>(npn, a->ndm);
      n_coarse = ngroups;
    }
    check(max_group <= 1024, "ndm_per_nominal must be <= 1024 on the GPU");
    check_overflow(ngroups <= 0xFFFFFFull && n_coarse <= 0xFFFFFFFFull,
                   "too many nominal DM groups");
    const uint32_t flush = static_cast<uint32_t>(
        std::max<uint64_t>(1, 65535 / (255 * a->subband_channels)));
    uint64_t lds = 0;
    const uint64_t halo_sum = subband_halo_sum(nsub, a->max_residual, a->residual_window);
    const uint32_t chunk = subband_chunk(nsub, max_group, halo_sum, lds);
    check_overflow(chunk != 0, "subband dedispersion delays too large for LDS");
    // spans: enough workgroups to fill the chip a few times over, each at
    // least 16 chunks long so the halo is a small part of its stage-1 work
    const uint64_t out_n = a->out_nsamples;
    const uint64_t max_spans = std::max<uint64_t>(1, ceil_div(out_n, 16ull * chunk));
    uint64_t nspans = std::min(max_spans, std::max<uint64_t>(1, ceil_div(2048, ngroups)));
    const uint64_t span = ceil_div(ceil_div(out_n, nspans), chunk) * chunk;
    nspans = ceil_div(out_n, span);
    check_overflow(span <= 0xFFFFFFFFull && ngroups * nspans < 0x7FFFFFF0ull,
                   "dedispersion grid too large");
    // every stage-1 read of the last span stays inside the padded rows
    check(ceil_div(out_n, chunk) * chunk +
                  ((static_cast<uint64_t>(a->max_residual) + 7) & ~7ull) + kSbReadPad +
                  static_cast<uint64_t>(a->max_coarse) <=
              row_len,
          "channel rows are not padded far enough for the delays (see "
          "gaffa_hip_subband_channel_stride)");
    const uint64_t nsub4 = (nsub + 3) & ~3ull;
    const uint64_t window_cap = (lds - 4 * (max_group * nsub4 + 2 * nsub4)) / 2;
    const uint64_t nwork = ngroups * nspans;
    const uint64_t grid = ceil_div(nwork, kXcds) * kXcds;  // whole rounds over the XCDs
    const bool wide = 2 * lds > kMaxLds;
    const void* kernel = wide
        ? reinterpret_cast<const void*>(dedisperse_subband_stream_kernel<kSbThreads>)
        : reinterpret_cast<const void*>(dedisperse_subband_stream_kernel<kSbThreadsSmall>);
    hip_check(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds)),
              "hipFuncSetAttribute");
    auto launch = [&](auto kern, uint32_t threads) {
      hipLaunchKernelGGL(kern, dim3(static_cast<uint32_t>(grid)),
                       dim3(threads), static_cast<uint32_t>(lds), as_stream(stream),
                       a->channels, a->channel_stride, row_len, static_cast<uint32_t>(nsel),
                       static_cast<uint32_t>(a->subband_channels), static_cast<uint32_t>(nsub),
                       reinterpret_cast<const SubbandGroupDev*>(a->groups),
                       static_cast<uint32_t>(ngroups), static_cast<uint32_t>(npn),
                       static_cast<uint32_t>(a->ndm), static_cast<uint32_t>(n_coarse),
                       a->coarse_delays, a->residual_delays, a->out, out_n, a->out_stride,
                       chunk, static_cast<uint32_t>(span), static_cast<uint32_t>(nspans),
                       static_cast<uint32_t>(max_group), static_cast<uint32_t>(window_cap),
                       flush, error_flag_for_device());
    };
    if (wide) launch(dedisperse_subband_stream_kernel<kSbThreads>, kSbThreads);
    else launch(dedisperse_subband_stream_kernel<kSbThreadsSmall>, kSbThreadsSmall);
    launch_check("dedisperse_subband_stream_kernel");
  });
}

uint32_t gaffa_hip_subband_time_tile(uint64_t nsel, uint64_t subband_channels,
                                     uint64_t max_group_ndm, int32_t max_residual,
                                     int64_t residual_window) {
  if (nsel == 0 || subband_channels == 0 || max_group_ndm == 0 || max_residual < 0) return 0;
  const uint64_t nsub = ceil_div(nsel, subband_channels);
  uint64_t lds = 0;
  return subband_chunk(nsub, max_group_ndm, subband_halo_sum(nsub, max_residual, residual_window),
                       lds);
}

uint64_t gaffa_hip_subband_channel_stride(uint64_t nsamples, int32_t max_coarse,
                                          int32_t max_residual) {
  // the last span's stage-1 passes read up to one chunk (<= 512) past the
  // output range, plus the residual halo, the pass padding and the coarse delay
  const uint64_t need = nsamples + static_cast<uint64_t>(std::max(max_coarse, 0)) +
                        static_cast<uint64_t>(std::max(max_residual, 0)) + 2048 + 64;
  return (need + 255) / 256 * 256;
}

int gaffa_hip_dedisperse_check(gaffa_hip_stream stream) {
  return guarded([&] { check_dedispersion_guard(as_stream(stream)); });
}

}  // extern "C"

// ------------------------------------------------- subband spectrum ----
// Subband sums of the aligned dynamic spectrum at one DM, for folding:
// out[k][t] = sum over the channels c of subband k of x[t + delay[c]][c],
// read from the channel-major copy.  gaffa folds the aligned spectrum
// (dedispersion.cpp:249-289) and averages each output channel's inputs in
// FP64 (folding.cpp:354-360); the sum of uint8 samples is an exact integer
// there, so folding these u16 sums divided by the channel count gives the
// same values.  One wave per (512-sample tile, subband), 8 samples per lane,
// with the stage-1 reads of subband dedispersion.
namespace gaffa_hip {
namespace {

constexpr uint32_t kSsTile = 512;

__global__ void __launch_bounds__(64)
subband_spectrum_kernel(const uint8_t* __restrict__ cm, uint64_t cm_stride, uint32_t nsel,
                        uint32_t sbc, const int32_t* __restrict__ delays,
                        uint16_t* __restrict__ out, uint64_t out_stride, uint64_t out_nsamples) {
  const uint32_t lane = threadIdx.x;
  const uint32_t k = blockIdx.y;
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kSsTile;
  const uint32_t cs = k * sbc;
  const uint32_t ce = min(cs + sbc, nsel);
  uint32_t a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  uint32_t c0 = cs;
  for (; c0 + kSbUnroll <= ce; c0 += kSbUnroll)
    sb_channels<kSbUnroll>(cm + static_cast<uint64_t>(c0) * cm_stride, cm_stride, delays + c0,
                           t0, lane, a0, a1, a2, a3);
  for (; c0 < ce; ++c0)
    sb_channels<1>(cm + static_cast<uint64_t>(c0) * cm_stride, cm_stride, delays + c0, t0, lane,
                   a0, a1, a2, a3);
  const uint64_t t = t0 + 8u * lane;
  if (t >= out_nsamples) return;
  uint16_t* o = out + static_cast<uint64_t>(k) * out_stride + t;
  uint4 v;
  v.x = __builtin_amdgcn_perm(a1, a0, 0x05040100u);  // s0 | s1 << 16
  v.y = __builtin_amdgcn_perm(a1, a0, 0x07060302u);  // s2 | s3 << 16
  v.z = __builtin_amdgcn_perm(a3, a2, 0x05040100u);
  v.w = __builtin_amdgcn_perm(a3, a2, 0x07060302u);
  if (t + 8 <= out_nsamples) {
    *reinterpret_cast<uint4*>(o) = v;
  } else {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
    for (uint32_t j = 0; t + j < out_nsamples; ++j)
      o[j] = static_cast<uint16_t>(w[j >> 1] >> (16u * (j & 1u)));
  }
}

}  // namespace
}  // namespace gaffa_hip

extern "C" int gaffa_hip_subband_spectrum_u8(const gaffa_hip_subband_spectrum_args* a,
                                             gaffa_hip_stream stream) {
  using namespace gaffa_hip;
  return guarded([&] {
    check(a != nullptr, "subband spectrum arguments must not be null");
    check(a->channels != nullptr && a->delays != nullptr && a->out != nullptr,
          "subband spectrum pointers must not be null");
    check(a->nsel > 0 && a->nsel <= 0xFFFFFFFFull, "channel count out of range");
    check(a->subband_channels > 0 && a->subband_channels <= 256,
          "subband spectrum supports 1..256 channels per subband");
    check(a->out_nsamples > 0, "subband spectrum output range is empty");
    check(a->max_delay >= 0, "subband spectrum delays must be non-negative");
    const uint64_t row_len = a->channel_length == 0 ? a->channel_stride : a->channel_length;
    check(row_len <= a->channel_stride || a->nsel == 1, "channel_length exceeds the row pitch");
    check(a->channel_stride % 4 == 0 && (reinterpret_cast<uintptr_t>(a->channels) & 3u) == 0,
          "channel rows must be dword aligned");
    const uint64_t nsub = ceil_div(a->nsel, a->subband_channels);
    check(nsub <= 65535, "too many subbands for one launch");
    check(a->out_stride >= a->out_nsamples && a->out_stride % 8 == 0 &&
              (reinterpret_cast<uintptr_t>(a->out) & 15u) == 0,
          "subband spectrum rows must be 16-byte aligned multiples of 8 samples");
    // every read of the last tile stays inside the rows (the kernel trusts
    // max_delay to bound the table)
    check(ceil_div(a->out_nsamples, kSsTile) * kSsTile + static_cast<uint64_t>(a->max_delay) +
                  kSbReadPad <=
              row_len,
          "channel rows are not padded far enough for the delays");
    const uint64_t tiles = ceil_div(a->out_nsamples, kSsTile);
    check_overflow(tiles < 0x7FFFFFFFull, "subband spectrum too long");
    hipLaunchKernelGGL(subband_spectrum_kernel,
                       dim3(static_cast<uint32_t>(tiles), static_cast<uint32_t>(nsub)), dim3(64),
                       0, as_stream(stream), a->channels, a->channel_stride,
                       static_cast<uint32_t>(a->nsel), static_cast<uint32_t>(a->subband_channels),
                       a->delays, a->out, a->out_stride, a->out_nsamples);
    launch_check("subband_spectrum_kernel");
  });
}

// ------------------------------------------------------------ spectrum ----
// Aligned dynamic spectrum (dedispersion.cpp:249-289): out[t][c] =
// x[t + delay[c]][chan_begin + c] for t < out_nsamples, any element size.  One
// workgroup per (time tile, channel tile): the tile's input rows
// [t0 + dmin, t0 + tile + dmax) x tile channels are staged in LDS with loads
// along channels, then every output row of the tile is written with stores
// along channels.
namespace gaffa_hip {
namespace {

template <typename T>
__global__ void __launch_bounds__(kThreads)
spectrum_kernel(const T* __restrict__ x, uint64_t nsamples, uint64_t nchans, uint64_t chan_begin,
                uint32_t nsel, const int32_t* __restrict__ delays, T* __restrict__ out,
                uint64_t out_nsamples, uint32_t tile, uint32_t ctile, uint32_t slab_rows,
                int32_t* __restrict__ error_flag) {
  extern __shared__ uint32_t dyn_lds[];
  T* slab = reinterpret_cast<T*>(dyn_lds);  // [slab_rows][ctile]
  __shared__ int32_t sdel[256];
  __shared__ int32_t smin, smax;
  const uint32_t tid = threadIdx.x;
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * tile;
  const uint32_t c0 = blockIdx.y * ctile;
  const uint32_t nc = min(ctile, nsel - c0);
  if (tid == 0) {
    smin = 0x7fffffff;
    smax = 0;
  }
  __syncthreads();
  if (tid < nc) {
    const int32_t d = delays[c0 + tid];
    sdel[tid] = d;
    atomicMin(&smin, d);
    atomicMax(&smax, d);
  }
  __syncthreads();
  const int32_t dmin = smin;
  const uint32_t rows = tile + static_cast<uint32_t>(smax - dmin);
  if (rows > slab_rows) {  // host/kernel table mismatch: refuse the block
    if (tid == 0) atomicExch(error_flag, 1);
    return;
  }
  for (uint32_t e = tid; e < rows * nc; e += kThreads) {
    const uint32_t r = e / nc, c = e - r * nc;
    const uint64_t t = t0 + static_cast<uint64_t>(dmin) + r;
    slab[r * ctile + c] = t < nsamples ? x[t * nchans + chan_begin + c0 + c] : T(0);
  }
  __syncthreads();
  const uint32_t tcount = static_cast<uint32_t>(min<uint64_t>(tile, out_nsamples - t0));
  for (uint32_t e = tid; e < tcount * nc; e += kThreads) {
    const uint32_t tl = e / nc, c = e - tl * nc;
    out[(t0 + tl) * nsel + c0 + c] =
        slab[(tl + static_cast<uint32_t>(sdel[c] - dmin)) * ctile + c];
  }
}

// Aligned spectrum staged through LDS in 32-byte channel tiles (32 u8, 16
// u16 or 8 32-bit channels): block (channel tile x, time tile y) loads the
// rows [t0 + dmin, t0 + tile + dmax) of its tile with 16-byte loads, then
// writes each output row's 32 bytes as eight words gathered from the rows at
// each channel's delay.  Channel tiles are the fastest grid index, so the
// blocks of one time tile run together and share the input rows in L2.
constexpr uint32_t kSpecTileRows = 1024;  // output rows per block

template <typename T>
__global__ void __launch_bounds__(kThreads)
spectrum_tiled_kernel(const T* __restrict__ x, uint64_t nsamples, uint64_t nchans,
                      uint64_t chan_begin, const int32_t* __restrict__ delays,
                      uint32_t* __restrict__ out, uint64_t out_nsamples, uint32_t out_words,
                      uint32_t slab_rows, int32_t* __restrict__ error_flag) {
  constexpr uint32_t CT = 32 / sizeof(T);  // channels per tile
  constexpr uint32_t E = 4 / sizeof(T);    // channels per word
  extern __shared__ uint4 slab4[];         // [slab_rows][2] x 16 bytes
  __shared__ int32_t sdel[CT];
  __shared__ int32_t dmin_s, dmax_s;
  const uint32_t c0 = blockIdx.x * CT;
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.y) * kSpecTileRows;
  if (threadIdx.x < CT) sdel[threadIdx.x] = delays[c0 + threadIdx.x];
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t lo = sdel[0], hi = sdel[0];
    for (uint32_t c = 1; c < CT; ++c) {
      lo = min(lo, sdel[c]);
      hi = max(hi, sdel[c]);
    }
    dmin_s = lo;
    dmax_s = hi;
  }
  __syncthreads();
  const int32_t dmin = dmin_s;
  const uint32_t rows = kSpecTileRows + static_cast<uint32_t>(dmax_s - dmin);
  if (rows > slab_rows) {  // host/kernel table mismatch: refuse the block
    if (threadIdx.x == 0) atomicExch(error_flag, 1);
    return;
  }
  const T* src = x + chan_begin + c0;
  for (uint32_t e = threadIdx.x; e < 2 * rows; e += kThreads) {
    const uint64_t t = t0 + static_cast<uint64_t>(dmin) + (e >> 1);
    uint4 v = make_uint4(0, 0, 0, 0);
    if (t < nsamples)
      v = *reinterpret_cast<const uint4*>(src + t * nchans + (e & 1u) * (CT / 2));
    slab4[e] = v;
  }
  __syncthreads();
  const T* slab = reinterpret_cast<const T*>(slab4);
  const uint32_t tcount = static_cast<uint32_t>(min<uint64_t>(kSpecTileRows, out_nsamples - t0));
  for (uint32_t e = threadIdx.x; e < tcount * 8u; e += kThreads) {
    const uint32_t tl = e >> 3, w = e & 7u;
    T v[E];
#pragma unroll
    for (uint32_t j = 0; j < E; ++j) {
      const uint32_t c = w * E + j;
      v[j] = slab[(tl + static_cast<uint32_t>(sdel[c] - dmin)) * CT + c];
    }
    uint32_t packed;
    __builtin_memcpy(&packed, v, 4);
    out[(t0 + tl) * out_words + blockIdx.x * 8u + w] = packed;
  }
}

// Aligned spectrum without staging, for channel counts that fill whole
// 32-bit words: one thread per output word (4 u8, 2 u16 or 1 32-bit
// channel) over a run of output rows; each element is read from its
// channel's delayed row (0 past the end, as gaffa's reader) and the word is
// stored whole.  Lanes of a wave take adjacent words, so a wave reads
// adjacent channels of nearly the same rows.
constexpr uint32_t kSpecRows = 64;  // output rows per block

template <typename T>
__global__ void __launch_bounds__(kThreads)
spectrum_direct_kernel(const T* __restrict__ x, uint64_t nsamples, uint64_t nchans,
                       uint64_t chan_begin, uint32_t nsel, const int32_t* __restrict__ delays,
                       uint32_t* __restrict__ out, uint64_t out_nsamples) {
  constexpr uint32_t E = 4 / sizeof(T);
  const uint32_t words = nsel / E;
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * kSpecRows;
  const uint64_t t1 = min(out_nsamples, t0 + kSpecRows);
  for (uint32_t w = threadIdx.x; w < words; w += kThreads) {
    uint64_t d[E];
#pragma unroll
    for (uint32_t j = 0; j < E; ++j) d[j] = static_cast<uint64_t>(delays[w * E + j]);
    const T* col = x + chan_begin + static_cast<uint64_t>(w) * E;
    uint64_t t = t0;
    for (; t + 4 <= t1; t += 4) {
      T v[4][E];
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u)
#pragma unroll
        for (uint32_t j = 0; j < E; ++j) {
          const uint64_t r = t + u + d[j];
          v[u][j] = r < nsamples ? col[min(r, nsamples - 1) * nchans + j] : T(0);
        }
#pragma unroll
      for (uint32_t u = 0; u < 4; ++u) {
        uint32_t packed;
        __builtin_memcpy(&packed, v[u], 4);
        out[(t + u) * words + w] = packed;
      }
    }
    for (; t < t1; ++t) {
      T v[E];
#pragma unroll
      for (uint32_t j = 0; j < E; ++j) {
        const uint64_t r = t + d[j];
        v[j] = r < nsamples ? col[min(r, nsamples - 1) * nchans + j] : T(0);
      }
      uint32_t packed;
      __builtin_memcpy(&packed, v, 4);
      out[t * words + w] = packed;
    }
  }
}

}  // namespace
}  // namespace gaffa_hip

extern "C" int gaffa_hip_dedisperse_spectrum(const gaffa_hip_spectrum_args* a,
                                             gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr && a->samples != nullptr && a->delays != nullptr && a->out != nullptr,
          "dedispersion device pointers must not be null");
    check(a->nsamples > 0 && a->nchans > 0, "dedispersion requires a non-empty filterbank");
    check(a->chan_begin < a->chan_end && a->chan_end <= a->nchans,
          "channel range exceeds sample channel count");
    check(a->out_nsamples > 0 && a->out_nsamples <= a->nsamples,
          "dedispersion valid output range is empty");
    check(a->element_size == 1 || a->element_size == 2 || a->element_size == 4,
          "spectrum element size must be 1, 2 or 4 bytes");
    check(a->max_tile_spread >= 0, "dedispersion requires non-negative delays");
    const uint64_t nsel = a->chan_end - a->chan_begin;
    check_overflow(nsel <= 0xFFFFFFFFull, "dedispersion channel range too large");
    const hipStream_t st = as_stream(stream);
    const uint32_t ct32 = 32u / a->element_size;  // channels per 32-byte tile
    if ((nsel * a->element_size) % 32 == 0 && (a->nchans * a->element_size) % 16 == 0 &&
        (a->chan_begin * a->element_size) % 16 == 0 &&
        (reinterpret_cast<uintptr_t>(a->samples) & 15u) == 0 &&
        (reinterpret_cast<uintptr_t>(a->out) & 3u) == 0 && a->channel_tile >= ct32 &&
        a->channel_tile % ct32 == 0) {
      // every 32-byte tile lies inside one channel_tile tile: its spread is bounded
      const uint64_t slab_rows = kSpecTileRows + static_cast<uint64_t>(a->max_tile_spread);
      const uint64_t lds = slab_rows * 32;
      if (lds <= 64 * 1024) {
        const uint64_t ttiles = ceil_div(a->out_nsamples, kSpecTileRows);
        check_overflow(ttiles < 65536, "spectrum grid too large");
        const dim3 grid(static_cast<uint32_t>(nsel / ct32), static_cast<uint32_t>(ttiles));
        auto* o = static_cast<uint32_t*>(a->out);
        const uint32_t words = static_cast<uint32_t>(nsel * a->element_size / 4);
        int32_t* flag = error_flag_for_device();
        switch (a->element_size) {
          case 1:
            hipLaunchKernelGGL(spectrum_tiled_kernel<uint8_t>, grid, dim3(kThreads), lds, st,
                               static_cast<const uint8_t*>(a->samples), a->nsamples, a->nchans,
                               a->chan_begin, a->delays, o, a->out_nsamples, words,
                               static_cast<uint32_t>(slab_rows), flag);
            break;
          case 2:
            hipLaunchKernelGGL(spectrum_tiled_kernel<uint16_t>, grid, dim3(kThreads), lds, st,
                               static_cast<const uint16_t*>(a->samples), a->nsamples, a->nchans,
                               a->chan_begin, a->delays, o, a->out_nsamples, words,
                               static_cast<uint32_t>(slab_rows), flag);
            break;
          default:
            hipLaunchKernelGGL(spectrum_tiled_kernel<uint32_t>, grid, dim3(kThreads), lds, st,
                               static_cast<const uint32_t*>(a->samples), a->nsamples, a->nchans,
                               a->chan_begin, a->delays, o, a->out_nsamples, words,
                               static_cast<uint32_t>(slab_rows), flag);
        }
        launch_check("spectrum_tiled_kernel");
        return;
      }
    }
    if ((nsel * a->element_size) % 4 == 0 && (reinterpret_cast<uintptr_t>(a->out) & 3u) == 0) {
      const dim3 grid(static_cast<uint32_t>(ceil_div(a->out_nsamples, kSpecRows)));
      check_overflow(ceil_div(a->out_nsamples, kSpecRows) < 0x7FFFFFFFull,
                     "spectrum grid too large");
      auto* o = static_cast<uint32_t*>(a->out);
      switch (a->element_size) {
        case 1:
          hipLaunchKernelGGL(spectrum_direct_kernel<uint8_t>, grid, dim3(kThreads), 0, st,
                             static_cast<const uint8_t*>(a->samples), a->nsamples, a->nchans,
                             a->chan_begin, static_cast<uint32_t>(nsel), a->delays, o,
                             a->out_nsamples);
          break;
        case 2:
          hipLaunchKernelGGL(spectrum_direct_kernel<uint16_t>, grid, dim3(kThreads), 0, st,
                             static_cast<const uint16_t*>(a->samples), a->nsamples, a->nchans,
                             a->chan_begin, static_cast<uint32_t>(nsel), a->delays, o,
                             a->out_nsamples);
          break;
        default:
          hipLaunchKernelGGL(spectrum_direct_kernel<uint32_t>, grid, dim3(kThreads), 0, st,
                             static_cast<const uint32_t*>(a->samples), a->nsamples, a->nchans,
                             a->chan_begin, static_cast<uint32_t>(nsel), a->delays, o,
                             a->out_nsamples);
      }
      launch_check("spectrum_direct_kernel");
      return;
    }
    // widest tile (time x channels) whose slab fits 96 KiB
    const uint32_t ctile = a->channel_tile;
    check(ctile == 8 || ctile == 16 || ctile == 32 || ctile == 64 || ctile == 128 || ctile == 256,
          "spectrum channel tile must be a power of two in [8, 256]");
    uint32_t tile = 0, slab_rows = 0;
    for (uint32_t cand = 1024; cand >= 16 && tile == 0; cand /= 2) {
      const uint64_t rows = cand + static_cast<uint64_t>(a->max_tile_spread);
      if (rows * ctile * a->element_size <= 96 * 1024) {
        tile = cand;
        slab_rows = static_cast<uint32_t>(rows);
      }
    }
    check_overflow(tile != 0, "spectrum delay spread too large for LDS; use a smaller channel tile");
    const uint64_t tiles = ceil_div(a->out_nsamples, tile);
    const uint64_t ctiles = ceil_div(nsel, ctile);
    check_overflow(tiles < 0x7FFFFFFFull && ctiles < 65536, "spectrum grid too large");
    const uint32_t lds = slab_rows * ctile * a->element_size;
    const dim3 grid(static_cast<uint32_t>(tiles), static_cast<uint32_t>(ctiles));
    int32_t* flag = error_flag_for_device();
    switch (a->element_size) {
      case 1:
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(spectrum_kernel<uint8_t>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds)),
                  "hipFuncSetAttribute");
        hipLaunchKernelGGL(spectrum_kernel<uint8_t>, grid, dim3(kThreads), lds, st,
                           static_cast<const uint8_t*>(a->samples), a->nsamples, a->nchans,
                           a->chan_begin, static_cast<uint32_t>(nsel), a->delays,
                           static_cast<uint8_t*>(a->out), a->out_nsamples, tile, ctile, slab_rows,
                           flag);
        break;
      case 2:
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(spectrum_kernel<uint16_t>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds)),
                  "hipFuncSetAttribute");
        hipLaunchKernelGGL(spectrum_kernel<uint16_t>, grid, dim3(kThreads), lds, st,
                           static_cast<const uint16_t*>(a->samples), a->nsamples, a->nchans,
                           a->chan_begin, static_cast<uint32_t>(nsel), a->delays,
                           static_cast<uint16_t*>(a->out), a->out_nsamples, tile, ctile,
                           slab_rows, flag);
        break;
      default:
        hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(spectrum_kernel<uint32_t>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds)),
                  "hipFuncSetAttribute");
        hipLaunchKernelGGL(spectrum_kernel<uint32_t>, grid, dim3(kThreads), lds, st,
                           static_cast<const uint32_t*>(a->samples), a->nsamples, a->nchans,
                           a->chan_begin, static_cast<uint32_t>(nsel), a->delays,
                           static_cast<uint32_t*>(a->out), a->out_nsamples, tile, ctile,
                           slab_rows, flag);
        break;
    }
    launch_check("spectrum_kernel");
  });
}

// ------------------------------------------------------------- row copy ----
namespace gaffa_hip {
namespace {

__global__ void __launch_bounds__(kThreads)
copy_rows_kernel(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst, uint64_t total,
                 uint64_t nchans, int32_t reverse) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * kThreads + threadIdx.x;
  if (e >= total) return;
  const uint64_t r = e / nchans, c = e - r * nchans;
  dst[e] = src[r * nchans + (reverse ? nchans - 1 - c : c)];
}

}  // namespace
}  // namespace gaffa_hip

extern "C" int gaffa_hip_copy_rows_u8(const uint8_t* src, uint8_t* dst, uint64_t rows,
                                      uint64_t nchans, int32_t reverse, gaffa_hip_stream stream) {
  return guarded([&] {
    check(src != nullptr && dst != nullptr, "row copy pointers must not be null");
    check(nchans > 0, "row copy requires channels");
    const uint64_t total = rows * nchans;
    if (total == 0) return;
    check_overflow(total / kThreads < (1ull << 31), "row copy too large");
    hipLaunchKernelGGL(copy_rows_kernel, dim3(static_cast<uint32_t>(ceil_div(total, kThreads))),
                       dim3(kThreads), 0, as_stream(stream), src, dst, total, nchans, reverse);
    launch_check("copy_rows_kernel");
  });
}
