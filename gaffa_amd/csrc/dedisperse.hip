// Incoherent dedispersion of uint8 filterbanks on gfx950.
//
// Both kernels stage delay-shifted input through LDS: a workgroup copies a
// time slab of a channel block from HBM (row segments of contiguous channel
// bytes, coalesced), transposes it to channel-major bytes in LDS, then every
// thread reads 4 consecutive time samples of a channel with one aligned dword
// pair + v_alignbyte and accumulates them as packed 16-bit pairs (<= 32 or
// subband_channels channels per flush, so no 16-bit lane can overflow).
// Integer sums are order-independent, so results equal gaffa's CPU sums bit
// for bit (dedispersion.cpp:197-215, 291-329, 351-453).
#include "common.hpp"

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

// ------------------------------------------------------------- subband ----
// One workgroup = (time tile of `tile` outputs, nominal DM group k).
// Stage 1 builds inter[s][j], j in [0, tile + max_residual + pad), as u16 in
// LDS, channel chunk by channel chunk through a byte slab; stage 2 sums the
// subbands at the residual delays for the group's DMs.
constexpr int kSbChunk = 32;       // channels per slab chunk
constexpr int kSbJ = 8;            // stage-1 samples per thread (jlen <= 2048)
constexpr uint32_t kSlabStride = 36;  // slab row stride in bytes (9 dwords)

__global__ void __launch_bounds__(kThreads)
dedisperse_subband_u8_kernel(const uint8_t* __restrict__ x, uint64_t nsamples,
                             uint64_t nchans, uint64_t chan_begin,
                             uint32_t nsel, uint32_t sbc, uint32_t nsub,
                             uint32_t npn, uint32_t ndm,
                             const int32_t* __restrict__ coarse,
                             const int32_t* __restrict__ resid,
                             uint32_t* __restrict__ out, uint64_t out_nsamples,
                             uint64_t out_stride, uint32_t tile,
                             uint32_t inter_len, uint32_t slab_len,
                             uint32_t flush_every,
                             int32_t* __restrict__ error_flag) {
  extern __shared__ uint32_t dyn_lds[];
  // [nsub][inter_len] u16, then [kSbChunk][slab_len] u8, then residuals
  uint16_t* inter = reinterpret_cast<uint16_t*>(dyn_lds);
  uint8_t* slab = reinterpret_cast<uint8_t*>(inter + static_cast<size_t>(nsub) * inter_len);
  int32_t* sres = reinterpret_cast<int32_t*>(slab + static_cast<size_t>(slab_len) * kSlabStride);
  __shared__ int32_t sdel[kSbChunk];
  __shared__ int32_t sminmax[2];

  const uint32_t tid = threadIdx.x;
  const uint64_t t0 = static_cast<uint64_t>(blockIdx.x) * tile;
  const uint32_t k = blockIdx.y;
  const uint32_t dm0 = k * npn;
  const uint32_t nd = min(npn, ndm - dm0);
  const uint32_t jlen = inter_len - 8;  // stage-1 samples computed
  const uint32_t nquads = jlen / 4;

  __shared__ int32_t sbad;
  if (tid == 0) sbad = 0;
  __syncthreads();
  // stage 2 reads inter[j + r .. j + r + 5] with j <= tile - 4: r must stay
  // within the staged halo or the block is refused (uniformly)
  const int32_t rmax = static_cast<int32_t>(jlen - tile - 4);
  for (uint32_t e = tid; e < nd * nsub; e += kThreads) {
    const int32_t r = resid[static_cast<uint64_t>(dm0) * nsub + e];
    if (r < 0 || r > rmax) sbad = 1;
    sres[e] = r;
  }
  __syncthreads();
  if (sbad) {
    if (tid == 0) atomicExch(error_flag, 1);
    return;
  }

  // Stage 1.  The slab is row-major [row][channel] with a 36-byte (9-dword,
  // odd) row stride: it is filled with 16-byte global loads, and lane j of a
  // wave reads byte (j + delay_c - dmin, c), so the lanes of one read hit
  // 64 distinct banks.  Thread tid owns samples j = tid + 256 q.
  (void)nquads;
  const bool vec_rows = (nchans % 16u) == 0 && (chan_begin % 16u) == 0 && (sbc % 32u) == 0 &&
                        (reinterpret_cast<uintptr_t>(x) & 15u) == 0;
  for (uint32_t s = 0; s < nsub; ++s) {
    const uint32_t cs = s * sbc;
    const uint32_t ce = min(cs + sbc, nsel);
    uint32_t acc[kSbJ];
#pragma unroll
    for (int q = 0; q < kSbJ; ++q) acc[q] = 0;

    for (uint32_t c0 = cs; c0 < ce; c0 += kSbChunk) {
      const uint32_t nc = min(static_cast<uint32_t>(kSbChunk), ce - c0);
      if (tid == 0) {
        sminmax[0] = 0x7fffffff;
        sminmax[1] = 0;
      }
      __syncthreads();
      if (tid < nc) {
        const int32_t d = coarse[static_cast<uint64_t>(k) * nsel + c0 + tid];
        sdel[tid] = d;
        atomicMin(&sminmax[0], d);
        atomicMax(&sminmax[1], d);
      }
      __syncthreads();
      const int32_t dmin = sminmax[0];
      const uint32_t rows = jlen + static_cast<uint32_t>(sminmax[1] - dmin) + 8;
      if (rows > slab_len) {
        if (tid == 0) atomicExch(error_flag, 1);
        return;
      }
      const uint64_t tbase = t0 + static_cast<uint64_t>(dmin);
      if (vec_rows && nc == kSbChunk) {
        uint32_t* slab32 = reinterpret_cast<uint32_t*>(slab);
        for (uint32_t e = tid; e < rows * 2; e += kThreads) {
          const uint32_t r = e >> 1, h = e & 1u;
          const uint64_t t = tbase + r;
          uint4 v = make_uint4(0, 0, 0, 0);
          if (t < nsamples)
            v = *reinterpret_cast<const uint4*>(x + t * nchans + chan_begin + c0 + 16u * h);
          uint32_t* dst = slab32 + (r * kSlabStride + 16u * h) / 4u;
          dst[0] = v.x;
          dst[1] = v.y;
          dst[2] = v.z;
          dst[3] = v.w;
        }
      } else {
        for (uint32_t e = tid; e < rows * nc; e += kThreads) {
          const uint32_t r = e / nc, c = e % nc;
          const uint64_t t = tbase + r;
          uint8_t v = 0;
          if (t < nsamples) v = x[t * nchans + chan_begin + c0 + c];
          slab[r * kSlabStride + c] = v;
        }
      }
      __syncthreads();
      for (uint32_t c = 0; c < nc; ++c) {
        const uint8_t* col = slab + static_cast<uint32_t>(sdel[c] - dmin) * kSlabStride + c;
#pragma unroll
        for (int q = 0; q < kSbJ; ++q) {
          const uint32_t j = tid + static_cast<uint32_t>(q) * kThreads;
          if (j < jlen) acc[q] += col[j * kSlabStride];
        }
      }
      __syncthreads();
    }
    uint16_t* row = inter + static_cast<size_t>(s) * inter_len;
#pragma unroll
    for (int q = 0; q < kSbJ; ++q) {
      const uint32_t j = tid + static_cast<uint32_t>(q) * kThreads;
      if (j < jlen) {
        row[j] = static_cast<uint16_t>(acc[q]);
      } else if (j < inter_len) {
        row[j] = 0;
      }
    }
  }
  __syncthreads();

  // Stage 2: thread -> (dm, quad) pairs.
  const uint32_t tquads = tile / 4;
  for (uint32_t e = tid; e < nd * tquads; e += kThreads) {
    const uint32_t i = e / tquads;
    const uint32_t q = e % tquads;
    uint32_t tot0 = 0, tot1 = 0, tot2 = 0, tot3 = 0;
    uint32_t p01 = 0, p23 = 0, pending = 0;
    for (uint32_t s = 0; s < nsub; ++s) {
      const uint32_t j = 4u * q + static_cast<uint32_t>(sres[i * nsub + s]);
      const uint32_t* row32 =
          reinterpret_cast<const uint32_t*>(inter + static_cast<size_t>(s) * inter_len);
      const uint32_t a = j >> 1;
      uint32_t w01, w23;
      if ((j & 1u) == 0) {
        w01 = row32[a];
        w23 = row32[a + 1];
      } else {
        const uint32_t d0 = row32[a], d1 = row32[a + 1], d2 = row32[a + 2];
        w01 = __builtin_amdgcn_alignbyte(d1, d0, 2);
        w23 = __builtin_amdgcn_alignbyte(d2, d1, 2);
      }
      p01 += w01;
      p23 += w23;
      if (++pending == flush_every) {
        tot0 += p01 & 0xFFFFu;
        tot1 += p01 >> 16;
        tot2 += p23 & 0xFFFFu;
        tot3 += p23 >> 16;
        p01 = p23 = 0;
        pending = 0;
      }
    }
    tot0 += p01 & 0xFFFFu;
    tot1 += p01 >> 16;
    tot2 += p23 & 0xFFFFu;
    tot3 += p23 >> 16;
    uint32_t* o = out + static_cast<uint64_t>(dm0 + i) * out_stride;
    const uint64_t t = t0 + 4u * q;
    if (t + 3 < out_nsamples && (out_stride & 3u) == 0 &&
        (reinterpret_cast<uintptr_t>(out) & 15u) == 0) {
      *reinterpret_cast<uint4*>(o + t) = make_uint4(tot0, tot1, tot2, tot3);
    } else {
      const uint32_t v[4] = {tot0, tot1, tot2, tot3};
#pragma unroll
      for (int jj = 0; jj < 4; ++jj)
        if (t + jj < out_nsamples) o[t + jj] = v[jj];
    }
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

int gaffa_hip_dedisperse_subband_u8(const gaffa_hip_dedisperse_subband_args* a,
                                    gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr, "dedispersion arguments must not be null");
    check(a->samples != nullptr && a->coarse_delays != nullptr &&
              a->residual_delays != nullptr && a->out != nullptr,
          "dedispersion device pointers must not be null");
    check(a->nsamples > 0, "dedispersion requires at least one sample");
    check(a->nchans > 0, "dedispersion requires at least one channel");
    check(a->chan_begin < a->chan_end, "channel range must be non-empty");
    check(a->chan_end <= a->nchans, "channel range exceeds sample channel count");
    check(a->ndm > 0, "dedispersion requires at least one DM");
    check(a->subband_channels > 0, "subband_channels must be positive");
    check(a->ndm_per_nominal > 0, "ndm_per_nominal must be positive");
    check(a->subband_channels <= 256,
          "GPU subband dedispersion supports at most 256 channels per subband");
    check(a->out_nsamples > 0 && a->out_nsamples <= a->nsamples,
          "dedispersion valid output range is empty");
    check(a->out_stride >= a->out_nsamples, "output stride must be >= out_nsamples");
    check(a->max_residual >= 0 && a->max_coarse_spread >= 0,
          "dedispersion requires non-negative delays");
    const uint64_t nsel = a->chan_end - a->chan_begin;
    const uint64_t nsub = ceil_div(nsel, a->subband_channels);
    const uint64_t nnom = ceil_div(a->ndm, a->ndm_per_nominal);
    check(a->ndm_per_nominal <= 1024, "ndm_per_nominal must be <= 1024 on the GPU");
    const uint32_t flush = static_cast<uint32_t>(
        std::max<uint64_t>(1, 65535 / (255 * a->subband_channels)));
    // Pick the largest time tile whose LDS footprint fits.
    uint32_t tile = 0, inter_len = 0, slab_len = 0;
    uint64_t lds = 0;
    const uint64_t res_bytes = a->ndm_per_nominal * nsub * 4;
    for (uint32_t cand = 2048; cand >= 64; cand /= 2) {
      const uint32_t jlen = cand + static_cast<uint32_t>(a->max_residual) + 4;
      const uint32_t jlen4 = (jlen + 3) & ~3u;
      if (jlen4 > static_cast<uint32_t>(kThreads * kSbJ)) continue;
      const uint32_t il = jlen4 + 8;
      const uint32_t sl =
          (jlen4 + static_cast<uint32_t>(a->max_coarse_spread) + 8 + 3) & ~3u;
      const uint64_t bytes = nsub * il * 2 + static_cast<uint64_t>(sl) * kSlabStride + res_bytes;
      if (bytes <= kMaxLds - 4096) {
        tile = cand;
        inter_len = il;
        slab_len = sl;
        lds = bytes;
        break;
      }
    }
    check_overflow(tile != 0, "subband dedispersion delays too large for LDS");
    const uint64_t tiles = ceil_div(a->out_nsamples, tile);
    check_overflow(tiles < 0x7FFFFFFFull && nnom < 65536, "dedispersion grid too large");
    hip_check(hipFuncSetAttribute(reinterpret_cast<const void*>(dedisperse_subband_u8_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds)),
              "hipFuncSetAttribute");
    const dim3 grid(static_cast<uint32_t>(tiles), static_cast<uint32_t>(nnom));
    hipLaunchKernelGGL(dedisperse_subband_u8_kernel, grid, dim3(kThreads),
                       static_cast<uint32_t>(lds), as_stream(stream), a->samples,
                       a->nsamples, a->nchans, a->chan_begin, static_cast<uint32_t>(nsel),
                       static_cast<uint32_t>(a->subband_channels),
                       static_cast<uint32_t>(nsub),
                       static_cast<uint32_t>(a->ndm_per_nominal),
                       static_cast<uint32_t>(a->ndm), a->coarse_delays,
                       a->residual_delays, a->out, a->out_nsamples, a->out_stride, tile,
                       inter_len, slab_len, flush, error_flag_for_device());
    launch_check("dedisperse_subband_u8_kernel");
  });
}

int gaffa_hip_dedisperse_check(gaffa_hip_stream stream) {
  return guarded([&] { check_dedispersion_guard(as_stream(stream)); });
}

}  // extern "C"
