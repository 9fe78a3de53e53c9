// Candidate scorer on gfx950: the dense work of the four circular 1D
// convolutions and the two fully connected layers, forward and backward, is
// one LDS-tiled GEMM on v_mfma_f32_16x16x32_bf16 (bfloat16 operands, FP32
// accumulation); the rest of training (im2col / col2im, pooling, loss, ReLU
// masks, bias sums, Adam) is elementwise.
//
// Activations are channels-last: a conv layer's input for batch row b is
// [64 phase][C] bfloat16, so the im2col row of (b, p) is three contiguous
// C-vectors (phases p - 1, p, p + 1 mod 64) and every forward GEMM operand
// is K-contiguous (16-byte loads straight into LDS).
#include "common.hpp"

#include <algorithm>
#include <cmath>

namespace gaffa_hip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr uint32_t kL = GAFFA_HIP_SCORER_PHASE_BINS;  // 64
constexpr uint32_t kSub = GAFFA_HIP_SCORER_SUBINTS;   // 32
constexpr uint32_t kPatch = GAFFA_HIP_SCORER_PATCH;   // 128
constexpr uint32_t kC = GAFFA_HIP_SCORER_CHANNELS;    // 64
constexpr uint32_t kH = GAFFA_HIP_SCORER_HIDDEN;      // 512
constexpr uint32_t kF = 2 * kC + kPatch;              // 256
constexpr uint32_t kCin[4] = {kSub, kC, kC, kC};
constexpr uint32_t kMaxSplits = 32;
constexpr uint64_t kPartialFloats = 1ull * kMaxSplits * kC * 3 * kC;
constexpr uint32_t kColChunks = 64;

// Round to nearest even (finite inputs).
__device__ __forceinline__ uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}

__device__ __forceinline__ float from_bf16(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// ------------------------------------------------------------------ GEMM ----
constexpr uint32_t kBM = 64, kBN = 64, kBK = 32, kLdsK = kBK + 8;

__device__ __forceinline__ void gemm_store(const gaffa_hip_gemm_args& g, uint32_t i, uint32_t j,
                                           float v) {
  if (g.bias != nullptr) v += g.bias[j];
  if (g.relu) v = fmaxf(v, 0.0f);
  const uint64_t o = static_cast<uint64_t>(i) * g.ldc + j;
  if (g.c32 != nullptr) g.c32[o] = v;
  if (g.c16 != nullptr) g.c16[o] = to_bf16(v);
}

// 64 x 64 output tile per workgroup, four waves of 32 x 32 (2 x 2 MFMA
// blocks).  Operand fragments, per lane l of a 16 x 16 x 32 block: A row
// l % 16, k = 8 (l / 16) .. + 7; B column l % 16, the same k; result rows
// 4 (l / 16) .. + 3 of column l % 16.  kAFast: A is K-contiguous (a_col 1,
// a_row and the base 16-byte aligned); kBFast: B is K-contiguous (b_row 1).
template <bool kAFast, bool kBFast>
__global__ void __launch_bounds__(256) gemm_kernel(gaffa_hip_gemm_args g, uint32_t kchunk) {
  __shared__ __attribute__((aligned(16))) uint16_t as[kBM * kLdsK];
  __shared__ __attribute__((aligned(16))) uint16_t bs[kBN * kLdsK];
  const uint32_t i0 = blockIdx.y * kBM, j0 = blockIdx.x * kBN;
  const uint32_t kb = blockIdx.z * kchunk;
  const uint32_t ke = min(g.k, kb + kchunk);
  const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
  const uint32_t wm = wave >> 1, wn = wave & 1u;
  floatx4 acc[2][2];
  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y) acc[x][y] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};

  for (uint32_t k0 = kb; k0 < ke; k0 += kBK) {
    if constexpr (kAFast) {
      const uint32_t i = t >> 2, kk = (t & 3u) * 8u;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (i0 + i < g.m && k0 + kk < ke)
        v = *reinterpret_cast<const uint4*>(g.a + static_cast<int64_t>(i0 + i) * g.a_row + k0 + kk);
      *reinterpret_cast<uint4*>(&as[i * kLdsK + kk]) = v;
    } else {
      const bool i_fast = g.a_row == 1;
      for (uint32_t q = 0; q < 8; ++q) {
        const uint32_t e = t + 256u * q;
        const uint32_t i = i_fast ? e % kBM : e / kBK;
        const uint32_t kk = i_fast ? e / kBM : e % kBK;
        uint16_t v = 0;
        if (i0 + i < g.m && k0 + kk < ke)
          v = g.a[static_cast<int64_t>(i0 + i) * g.a_row + static_cast<int64_t>(k0 + kk) * g.a_col];
        as[i * kLdsK + kk] = v;
      }
    }
    if constexpr (kBFast) {
      const uint32_t j = t >> 2, kk = (t & 3u) * 8u;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (j0 + j < g.n && k0 + kk < ke)
        v = *reinterpret_cast<const uint4*>(g.b + static_cast<int64_t>(j0 + j) * g.b_col + k0 + kk);
      *reinterpret_cast<uint4*>(&bs[j * kLdsK + kk]) = v;
    } else {
      const bool j_fast = g.b_col == 1;
      for (uint32_t q = 0; q < 8; ++q) {
        const uint32_t e = t + 256u * q;
        const uint32_t j = j_fast ? e % kBN : e / kBK;
        const uint32_t kk = j_fast ? e / kBN : e % kBK;
        uint16_t v = 0;
        if (j0 + j < g.n && k0 + kk < ke)
          v = g.b[static_cast<int64_t>(k0 + kk) * g.b_row + static_cast<int64_t>(j0 + j) * g.b_col];
        bs[j * kLdsK + kk] = v;
      }
    }
    __syncthreads();
    bf16x8 af[2], bf[2];
    for (int x = 0; x < 2; ++x)
      af[x] = *reinterpret_cast<const bf16x8*>(
          &as[(wm * 32u + x * 16u + (lane & 15u)) * kLdsK + 8u * (lane >> 4)]);
    for (int y = 0; y < 2; ++y)
      bf[y] = *reinterpret_cast<const bf16x8*>(
          &bs[(wn * 32u + y * 16u + (lane & 15u)) * kLdsK + 8u * (lane >> 4)]);
    for (int x = 0; x < 2; ++x)
      for (int y = 0; y < 2; ++y)
        acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[x], bf[y], acc[x][y], 0, 0, 0);
    __syncthreads();
  }

  for (int x = 0; x < 2; ++x)
    for (int y = 0; y < 2; ++y)
      for (int r = 0; r < 4; ++r) {
        const uint32_t i = i0 + wm * 32u + x * 16u + 4u * (lane >> 4) + r;
        const uint32_t j = j0 + wn * 32u + y * 16u + (lane & 15u);
        if (i >= g.m || j >= g.n) continue;
        if (gridDim.z > 1)
          g.partial[(static_cast<uint64_t>(blockIdx.z) * g.m + i) * g.n + j] = acc[x][y][r];
        else
          gemm_store(g, i, j, acc[x][y][r]);
      }
}

__global__ void __launch_bounds__(256) split_reduce_kernel(gaffa_hip_gemm_args g) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  const uint64_t mn = static_cast<uint64_t>(g.m) * g.n;
  if (e >= mn) return;
  float v = 0.0f;
  for (uint32_t z = 0; z < g.k_splits; ++z) v += g.partial[z * mn + e];
  gemm_store(g, static_cast<uint32_t>(e / g.n), static_cast<uint32_t>(e % g.n), v);
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

void validate_gemm(const gaffa_hip_gemm_args& g) {
  check(g.a != nullptr && g.b != nullptr, "GEMM operands must not be null");
  check(g.m > 0 && g.n > 0 && g.k > 0, "GEMM sizes must be positive");
  check(g.c32 != nullptr || g.c16 != nullptr, "GEMM needs an output");
  check(g.ldc >= g.n, "GEMM ldc must be >= n");
  check(g.k_splits >= 1 && g.k_splits <= kMaxSplits, "GEMM k_splits must be in 1..32");
  check(ceil_div(g.m, kBM) < 65536 && ceil_div(g.n, kBN) < 65536, "GEMM grid too large");
  check(g.k_splits == 1 || g.partial != nullptr, "GEMM split partials must not be null");
}

void launch_gemm(const gaffa_hip_gemm_args& g, hipStream_t st) {
  validate_gemm(g);
  const uint32_t kchunk =
      static_cast<uint32_t>(ceil_div(ceil_div(g.k, g.k_splits), kBK) * kBK);
  const uint32_t splits = static_cast<uint32_t>(ceil_div(g.k, kchunk));
  gaffa_hip_gemm_args a = g;
  a.k_splits = splits;
  const bool af = g.a_col == 1 && g.a_row % 8 == 0 && g.k % 8 == 0 && aligned16(g.a);
  const bool bf = g.b_row == 1 && g.b_col % 8 == 0 && g.k % 8 == 0 && aligned16(g.b);
  const dim3 grid(static_cast<uint32_t>(ceil_div(g.n, kBN)),
                  static_cast<uint32_t>(ceil_div(g.m, kBM)), splits);
  if (af && bf)
    hipLaunchKernelGGL((gemm_kernel<true, true>), grid, dim3(256), 0, st, a, kchunk);
  else if (af)
    hipLaunchKernelGGL((gemm_kernel<true, false>), grid, dim3(256), 0, st, a, kchunk);
  else if (bf)
    hipLaunchKernelGGL((gemm_kernel<false, true>), grid, dim3(256), 0, st, a, kchunk);
  else
    hipLaunchKernelGGL((gemm_kernel<false, false>), grid, dim3(256), 0, st, a, kchunk);
  launch_check("gemm_kernel");
  if (splits > 1) {
    const uint64_t mn = static_cast<uint64_t>(g.m) * g.n;
    hipLaunchKernelGGL(split_reduce_kernel, dim3(static_cast<uint32_t>(ceil_div(mn, 256))),
                       dim3(256), 0, st, a);
    launch_check("split_reduce_kernel");
  }
}

gaffa_hip_gemm_args gemm(const uint16_t* a, int64_t a_row, int64_t a_col, const uint16_t* b,
                         int64_t b_row, int64_t b_col, uint32_t m, uint32_t n, uint32_t k) {
  gaffa_hip_gemm_args g{};
  g.a = a;
  g.a_row = a_row;
  g.a_col = a_col;
  g.b = b;
  g.b_row = b_row;
  g.b_col = b_col;
  g.m = m;
  g.n = n;
  g.k = k;
  g.ldc = n;
  g.k_splits = 1;
  return g;
}

// ------------------------------------------------------- network layout ----
gaffa_hip_scorer_layout make_layout() {
  gaffa_hip_scorer_layout l{};
  const uint64_t rows[6] = {kC, kC, kC, kC, kH, 1};
  const uint64_t cols[6] = {3 * kCin[0], 3 * kCin[1], 3 * kCin[2], 3 * kCin[3], kF, kH};
  uint64_t o = 0;
  for (int i = 0; i < 6; ++i) {
    l.weight[i] = o;
    o += rows[i] * cols[i];
    l.bias[i] = o;
    o += rows[i];
    l.fan_in[i] = cols[i];
  }
  l.nparams = (o + 7) / 8 * 8;
  return l;
}

// Workspace carve-up for batch capacity B (R = 64 B rows of conv activations).
struct Work {
  uint16_t* x0;        // [R][32]
  uint16_t* xcol[4];   // [R][3 Cin]
  uint16_t* y[4];      // [R][64]
  uint16_t* f;         // [B][256]
  uint8_t* argmax;     // [B][64]
  uint16_t* h;         // [B][512]
  float* z;            // [B]
  float* dz;           // [B]
  uint16_t* dz16;      // [B] (+ pad)
  float* dh;           // [B][512]
  uint16_t* dh16;      // [B][512]
  float* df;           // [B][128]
  float* dy;           // [R][64]
  uint16_t* dy16;      // [R][64]
  float* dxcol;        // [R][192]
  float* partial;      // [kPartialFloats]
  float* colpart;      // [kColChunks][512]
  uint64_t bytes;
};

Work carve(void* base, uint32_t cap) {
  Work w{};
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    char* p = base != nullptr ? static_cast<char*>(base) + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return p;
  };
  const uint64_t B = cap, R = 64ull * cap;
  w.x0 = reinterpret_cast<uint16_t*>(take(R * kSub * 2));
  for (int l = 0; l < 4; ++l) {
    w.xcol[l] = reinterpret_cast<uint16_t*>(take(R * 3 * kCin[l] * 2));
    w.y[l] = reinterpret_cast<uint16_t*>(take(R * kC * 2));
  }
  w.f = reinterpret_cast<uint16_t*>(take(B * kF * 2));
  w.argmax = reinterpret_cast<uint8_t*>(take(B * kC));
  w.h = reinterpret_cast<uint16_t*>(take(B * kH * 2));
  w.z = reinterpret_cast<float*>(take(B * 4));
  w.dz = reinterpret_cast<float*>(take(B * 4));
  w.dz16 = reinterpret_cast<uint16_t*>(take(B * 2 + 16));
  w.dh = reinterpret_cast<float*>(take(B * kH * 4));
  w.dh16 = reinterpret_cast<uint16_t*>(take(B * kH * 2));
  w.df = reinterpret_cast<float*>(take(B * 2 * kC * 4));
  w.dy = reinterpret_cast<float*>(take(R * kC * 4));
  w.dy16 = reinterpret_cast<uint16_t*>(take(R * kC * 2));
  w.dxcol = reinterpret_cast<float*>(take(R * 3 * kC * 4));
  w.partial = reinterpret_cast<float*>(take(kPartialFloats * 4));
  w.colpart = reinterpret_cast<float*>(take(1ull * kColChunks * kH * 4));
  w.bytes = off;
  return w;
}

// ------------------------------------------------------ forward kernels ----
__device__ __forceinline__ uint32_t row_of(const uint32_t* index, uint32_t b) {
  return index != nullptr ? index[b] : b;
}

// x0[b][p][s] = bf16(cube[idx b][s][p]); f[b][128 + q] = bf16(patch[idx b][q]).
__global__ void __launch_bounds__(256)
gather_inputs_kernel(const float* __restrict__ cubes, const float* __restrict__ patches,
                     const uint32_t* __restrict__ index, uint32_t batch, uint16_t* __restrict__ x0,
                     uint16_t* __restrict__ f) {
  const uint32_t b = blockIdx.x;
  if (b >= batch) return;
  const uint64_t r = row_of(index, b);
  const float* cube = cubes + r * (kSub * kL);
  for (uint32_t e = threadIdx.x; e < kSub * kL; e += 256) {
    const uint32_t p = e / kSub, s = e % kSub;
    x0[(static_cast<uint64_t>(b) * kL + p) * kSub + s] = to_bf16(cube[s * kL + p]);
  }
  if (threadIdx.x < kPatch)
    f[static_cast<uint64_t>(b) * kF + 2 * kC + threadIdx.x] =
        to_bf16(patches[r * kPatch + threadIdx.x]);
}

// xcol[b, p][tap * cin + c] = x[b, (p + tap - 1) mod 64][c], 8 channels per thread.
__global__ void __launch_bounds__(256)
im2col_kernel(const uint16_t* __restrict__ x, uint16_t* __restrict__ xcol, uint32_t cin,
              uint64_t rows) {
  const uint32_t vec = cin / 8;
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= rows * 3 * vec) return;
  const uint32_t c8 = static_cast<uint32_t>(e % vec);
  const uint32_t tap = static_cast<uint32_t>((e / vec) % 3);
  const uint64_t n = e / (3ull * vec);
  const uint64_t b = n / kL;
  const uint32_t p = static_cast<uint32_t>(n % kL);
  const uint32_t src = (p + tap + kL - 1) % kL;
  const uint4 v = *reinterpret_cast<const uint4*>(x + (b * kL + src) * cin + c8 * 8);
  *reinterpret_cast<uint4*>(xcol + n * 3 * cin + tap * cin + c8 * 8) = v;
}

// f[b][c] = mean over phase, f[b][64 + c] = max over phase (first phase on ties).
__global__ void __launch_bounds__(256)
pool_kernel(const uint16_t* __restrict__ y, uint32_t batch, uint16_t* __restrict__ f,
            uint8_t* __restrict__ argmax) {
  const uint32_t e = blockIdx.x * 256 + threadIdx.x;
  if (e >= batch * kC) return;
  const uint32_t b = e / kC, c = e % kC;
  const uint16_t* col = y + static_cast<uint64_t>(b) * kL * kC + c;
  float sum = 0.0f, best = from_bf16(col[0]);
  uint32_t arg = 0;
  for (uint32_t p = 0; p < kL; ++p) {
    const float v = from_bf16(col[p * kC]);
    sum += v;
    if (v > best) {
      best = v;
      arg = p;
    }
  }
  f[static_cast<uint64_t>(b) * kF + c] = to_bf16(sum * (1.0f / kL));
  f[static_cast<uint64_t>(b) * kF + kC + c] = to_bf16(best);
  argmax[e] = static_cast<uint8_t>(arg);
}

// ----------------------------------------------------- backward kernels ----
// Binary cross-entropy with logits, mean over the batch: loss, dz = (sigmoid(z)
// - y) / B (FP32 and bfloat16) and the fc2 bias gradient sum(dz).  One block.
__global__ void __launch_bounds__(256)
bce_kernel(const float* __restrict__ z, const float* __restrict__ labels,
           const uint32_t* __restrict__ index, uint32_t batch, float* __restrict__ dz,
           uint16_t* __restrict__ dz16, float* __restrict__ loss, float* __restrict__ dbias) {
  __shared__ float sl[256], sg[256];
  float l = 0.0f, g = 0.0f;
  const float inv = 1.0f / static_cast<float>(batch);
  for (uint32_t b = threadIdx.x; b < batch; b += 256) {
    const float x = z[b], y = labels[row_of(index, b)];
    l += fmaxf(x, 0.0f) - x * y + log1pf(expf(-fabsf(x)));
    const float d = (1.0f / (1.0f + expf(-x)) - y) * inv;
    dz[b] = d;
    dz16[b] = to_bf16(d);
    g += d;
  }
  sl[threadIdx.x] = l;
  sg[threadIdx.x] = g;
  __syncthreads();
  for (uint32_t s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      sl[threadIdx.x] += sl[threadIdx.x + s];
      sg[threadIdx.x] += sg[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *loss = sl[0] * inv;
    *dbias = sg[0];
  }
  // zero the bfloat16 padding past the batch (read by the k-padded GEMM)
  if (threadIdx.x < 8) dz16[batch + threadIdx.x] = 0;
}

// dh[b][j] = (h > 0) ? dz[b] * w2[j] : 0, both factors bfloat16 (the product
// is exact in FP32).
__global__ void __launch_bounds__(256)
fc1_grad_kernel(const uint16_t* __restrict__ h, const uint16_t* __restrict__ dz16,
                const uint16_t* __restrict__ w2, uint32_t batch, float* __restrict__ dh,
                uint16_t* __restrict__ dh16) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= static_cast<uint64_t>(batch) * kH) return;
  const uint32_t b = static_cast<uint32_t>(e / kH), j = static_cast<uint32_t>(e % kH);
  const float v = from_bf16(h[e]) > 0.0f ? from_bf16(dz16[b]) * from_bf16(w2[j]) : 0.0f;
  dh[e] = v;
  dh16[e] = to_bf16(v);
}

// Gradient into conv4's output from the pool: mean part df[b][c] / 64, max
// part df[b][64 + c] at the argmax phase; masked by ReLU (y > 0).
__global__ void __launch_bounds__(256)
pool_grad_kernel(const float* __restrict__ df, const uint8_t* __restrict__ argmax,
                 const uint16_t* __restrict__ y, uint32_t batch, float* __restrict__ dy,
                 uint16_t* __restrict__ dy16) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= static_cast<uint64_t>(batch) * kL * kC) return;
  const uint32_t c = static_cast<uint32_t>(e % kC);
  const uint32_t p = static_cast<uint32_t>((e / kC) % kL);
  const uint64_t b = e / (kL * kC);
  float g = df[b * 2 * kC + c] * (1.0f / kL);
  if (argmax[b * kC + c] == p) g += df[b * 2 * kC + kC + c];
  if (!(from_bf16(y[e]) > 0.0f)) g = 0.0f;
  dy[e] = g;
  dy16[e] = to_bf16(g);
}

// dx[b, p][c] = sum over taps t = 0, 1, 2 of dxcol[b, (p - t + 1) mod 64][t cin + c],
// masked by the previous layer's ReLU.
__global__ void __launch_bounds__(256)
col2im_kernel(const float* __restrict__ dxcol, const uint16_t* __restrict__ yprev, uint32_t cin,
              uint64_t rows, float* __restrict__ dy, uint16_t* __restrict__ dy16) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= rows * cin) return;
  const uint32_t c = static_cast<uint32_t>(e % cin);
  const uint64_t n = e / cin;
  const uint64_t b = n / kL;
  const uint32_t p = static_cast<uint32_t>(n % kL);
  float g = 0.0f;
  for (uint32_t t = 0; t < 3; ++t) {
    const uint64_t src = b * kL + (p + kL + 1 - t) % kL;
    g += dxcol[src * 3 * cin + t * cin + c];
  }
  if (!(from_bf16(yprev[e]) > 0.0f)) g = 0.0f;
  dy[e] = g;
  dy16[e] = to_bf16(g);
}

// Column sums of x [rows][cols] (FP32) in two passes: chunk partials, then
// the chunks in order.
__global__ void __launch_bounds__(256)
colsum_partial_kernel(const float* __restrict__ x, uint64_t rows, uint32_t cols,
                      float* __restrict__ part) {
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c >= cols) return;
  const uint64_t per = ceil_div(rows, kColChunks);
  const uint64_t r0 = blockIdx.y * per, r1 = min(rows, r0 + per);
  float s = 0.0f;
  for (uint64_t r = r0; r < r1; ++r) s += x[r * cols + c];
  part[static_cast<uint64_t>(blockIdx.y) * cols + c] = s;
}

__global__ void __launch_bounds__(256)
colsum_final_kernel(const float* __restrict__ part, uint32_t cols, float* __restrict__ out) {
  const uint32_t c = blockIdx.x * 256 + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  for (uint32_t k = 0; k < kColChunks; ++k) s += part[static_cast<uint64_t>(k) * cols + c];
  out[c] = s;
}

__global__ void __launch_bounds__(256)
adam_kernel(float* __restrict__ p, uint16_t* __restrict__ p16, const float* __restrict__ g,
            float* __restrict__ m, float* __restrict__ v, uint64_t n, float lr, float b1, float b2,
            float eps, float c1, float c2) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e >= n) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.0f - b1) * gi;
  const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  const float pi = p[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  p[e] = pi;
  p16[e] = to_bf16(pi);
}

__global__ void __launch_bounds__(256)
sync_kernel(const float* __restrict__ p, uint16_t* __restrict__ p16, uint64_t n) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (e < n) p16[e] = to_bf16(p[e]);
}

// ------------------------------------------------------ input features ----
__global__ void __launch_bounds__(kPatch)
gather_patch_kernel(gaffa_hip_scorer_patch_args a) {
  const uint32_t i = blockIdx.x, q = threadIdx.x;
  const int64_t r = static_cast<int64_t>(a.row[i]) - static_cast<int64_t>(kPatch / 2) + q;
  const uint64_t rc = static_cast<uint64_t>(
      min<int64_t>(max<int64_t>(r, 0), static_cast<int64_t>(a.rows) - 1));
  a.out[static_cast<uint64_t>(i) * kPatch + q] =
      a.periodogram[a.series[i] * a.series_stride + rc * a.widths + a.width[i]];
}

// One block per row: FP32 mean, then FP32 variance about it (fixed-order
// reductions), then (x - mean) / sqrt(variance), or only x - mean when
// `scale` is 0 (periodogram patches keep their S/N units).
__global__ void __launch_bounds__(256)
normalise_rows_kernel(float* __restrict__ x, uint32_t length, uint32_t scale) {
  __shared__ float red[256];
  float* row = x + static_cast<uint64_t>(blockIdx.x) * length;
  auto reduce = [&](float v) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    const float r = red[0];
    __syncthreads();
    return r;
  };
  float s = 0.0f;
  for (uint32_t k = threadIdx.x; k < length; k += 256) s += row[k];
  const float mean = reduce(s) / static_cast<float>(length);
  float v = 0.0f;
  for (uint32_t k = threadIdx.x; k < length; k += 256) {
    const float d = row[k] - mean;
    v += d * d;
  }
  const float var = reduce(v) / static_cast<float>(length);
  const float inv = !scale ? 1.0f : (var > 0.0f ? 1.0f / sqrtf(var) : 0.0f);
  for (uint32_t k = threadIdx.x; k < length; k += 256) row[k] = (row[k] - mean) * inv;
}

dim3 blocks(uint64_t n) { return dim3(static_cast<uint32_t>(ceil_div(n, 256))); }

void colsum(const float* x, uint64_t rows, uint32_t cols, float* part, float* out,
            hipStream_t st) {
  hipLaunchKernelGGL(colsum_partial_kernel, dim3(static_cast<uint32_t>(ceil_div(cols, 256)),
                                                 kColChunks),
                     dim3(256), 0, st, x, rows, cols, part);
  hipLaunchKernelGGL(colsum_final_kernel, dim3(static_cast<uint32_t>(ceil_div(cols, 256))),
                     dim3(256), 0, st, part, cols, out);
  launch_check("colsum_kernel");
}

void validate_net(const gaffa_hip_scorer_net* net, uint32_t batch) {
  check(net != nullptr && net->params != nullptr && net->params16 != nullptr &&
            net->workspace != nullptr,
        "scorer network buffers must not be null");
  check(batch > 0 && batch <= net->capacity, "scorer batch must be in 1..capacity");
  check(net->capacity <= 65535, "scorer capacity must be <= 65535");
}

// Splits along k for a weight-gradient GEMM: about 1024 rows of the batch
// each, within the partial-sum buffer.
uint32_t splits_for(uint64_t k, uint64_t mn) {
  const uint64_t fit = std::max<uint64_t>(1, kPartialFloats / mn);
  return static_cast<uint32_t>(
      std::min<uint64_t>({kMaxSplits, fit, std::max<uint64_t>(1, k / 1024)}));
}

}  // namespace
}  // namespace gaffa_hip

using namespace gaffa_hip;

extern "C" {

int gaffa_hip_gemm_bf16(const gaffa_hip_gemm_args* args, gaffa_hip_stream stream) {
  return guarded([&] {
    check(args != nullptr, "GEMM arguments must not be null");
    launch_gemm(*args, as_stream(stream));
  });
}

int gaffa_hip_scorer_get_layout(gaffa_hip_scorer_layout* layout) {
  return guarded([&] {
    check(layout != nullptr, "scorer layout must not be null");
    *layout = make_layout();
  });
}

uint64_t gaffa_hip_scorer_workspace_bytes(uint32_t batch) {
  return carve(nullptr, batch).bytes;
}

int gaffa_hip_scorer_forward(const gaffa_hip_scorer_net* net, const float* cubes,
                             const float* patches, const uint32_t* index, uint32_t batch,
                             float* logits, gaffa_hip_stream stream) {
  return guarded([&] {
    validate_net(net, batch);
    check(cubes != nullptr && patches != nullptr && logits != nullptr,
          "scorer inputs and logits must not be null");
    const hipStream_t st = as_stream(stream);
    const gaffa_hip_scorer_layout lay = make_layout();
    const Work w = carve(net->workspace, net->capacity);
    const uint64_t R = 64ull * batch;
    hipLaunchKernelGGL(gather_inputs_kernel, dim3(batch), dim3(256), 0, st, cubes, patches,
                       index, batch, w.x0, w.f);
    launch_check("gather_inputs_kernel");
    const uint16_t* x = w.x0;
    for (int l = 0; l < 4; ++l) {
      const uint32_t cin = kCin[l];
      hipLaunchKernelGGL(im2col_kernel, blocks(R * 3 * (cin / 8)), dim3(256), 0, st, x,
                         w.xcol[l], cin, R);
      launch_check("im2col_kernel");
      gaffa_hip_gemm_args g = gemm(w.xcol[l], 3 * cin, 1, net->params16 + lay.weight[l], 1,
                                   3 * cin, static_cast<uint32_t>(R), kC, 3 * cin);
      g.bias = net->params + lay.bias[l];
      g.relu = 1;
      g.c16 = w.y[l];
      launch_gemm(g, st);
      x = w.y[l];
    }
    hipLaunchKernelGGL(pool_kernel, blocks(1ull * batch * kC), dim3(256), 0, st, w.y[3], batch,
                       w.f, w.argmax);
    launch_check("pool_kernel");
    gaffa_hip_gemm_args g1 =
        gemm(w.f, kF, 1, net->params16 + lay.weight[4], 1, kF, batch, kH, kF);
    g1.bias = net->params + lay.bias[4];
    g1.relu = 1;
    g1.c16 = w.h;
    launch_gemm(g1, st);
    gaffa_hip_gemm_args g2 = gemm(w.h, kH, 1, net->params16 + lay.weight[5], 1, 0, batch, 1, kH);
    g2.bias = net->params + lay.bias[5];
    g2.c32 = w.z;
    launch_gemm(g2, st);
    hip_check(hipMemcpyAsync(logits, w.z, 4ull * batch, hipMemcpyDeviceToDevice, st),
              "hipMemcpyAsync");
  });
}

int gaffa_hip_scorer_backward(const gaffa_hip_scorer_net* net, const float* labels,
                              const uint32_t* index, uint32_t batch, float* loss,
                              gaffa_hip_stream stream) {
  return guarded([&] {
    validate_net(net, batch);
    check(net->grads != nullptr && labels != nullptr && loss != nullptr,
          "scorer gradients, labels and loss must not be null");
    const hipStream_t st = as_stream(stream);
    const gaffa_hip_scorer_layout lay = make_layout();
    const Work w = carve(net->workspace, net->capacity);
    const uint64_t R = 64ull * batch;
    float* grads = net->grads;
    const uint16_t* p16 = net->params16;

    hipLaunchKernelGGL(bce_kernel, dim3(1), dim3(256), 0, st, w.z, labels, index, batch, w.dz,
                       w.dz16, loss, grads + lay.bias[5]);
    launch_check("bce_kernel");
    // fc2 weights: dW[j] = sum_b dz[b] h[b][j]
    gaffa_hip_gemm_args gw2 = gemm(w.dz16, 0, 1, w.h, kH, 1, 1, kH, batch);
    gw2.c32 = grads + lay.weight[5];
    gw2.k_splits = splits_for(batch, kH);
    gw2.partial = w.partial;
    launch_gemm(gw2, st);
    // into fc1's output, through its ReLU
    hipLaunchKernelGGL(fc1_grad_kernel, blocks(1ull * batch * kH), dim3(256), 0, st, w.h, w.dz16,
                       p16 + lay.weight[5], batch, w.dh, w.dh16);
    launch_check("fc1_grad_kernel");
    colsum(w.dh, batch, kH, w.colpart, grads + lay.bias[4], st);
    // fc1 weights: dW[j][k] = sum_b dh[b][j] f[b][k]
    gaffa_hip_gemm_args gw1 = gemm(w.dh16, 1, kH, w.f, kF, 1, kH, kF, batch);
    gw1.c32 = grads + lay.weight[4];
    gw1.k_splits = splits_for(batch, 1ull * kH * kF);
    gw1.partial = w.partial;
    launch_gemm(gw1, st);
    // into the pooled features (the conv half of f only)
    gaffa_hip_gemm_args gf = gemm(w.dh16, kH, 1, p16 + lay.weight[4], kF, 1, batch, 2 * kC, kH);
    gf.c32 = w.df;
    launch_gemm(gf, st);
    hipLaunchKernelGGL(pool_grad_kernel, blocks(R * kC), dim3(256), 0, st, w.df, w.argmax,
                       w.y[3], batch, w.dy, w.dy16);
    launch_check("pool_grad_kernel");
    for (int l = 3; l >= 0; --l) {
      const uint32_t cin = kCin[l];
      // weights: dW[co][k] = sum_n dy[n][co] xcol[n][k]
      gaffa_hip_gemm_args gw = gemm(w.dy16, 1, kC, w.xcol[l], 3 * cin, 1, kC, 3 * cin,
                                    static_cast<uint32_t>(R));
      gw.c32 = grads + lay.weight[l];
      gw.k_splits = splits_for(R, 3ull * kC * cin);
      gw.partial = w.partial;
      launch_gemm(gw, st);
      colsum(w.dy, R, kC, w.colpart, grads + lay.bias[l], st);
      if (l == 0) break;
      // into the layer input: dxcol[n][k] = sum_co dy[n][co] W[co][k], then col2im
      gaffa_hip_gemm_args gx = gemm(w.dy16, kC, 1, p16 + lay.weight[l], 3 * cin, 1,
                                    static_cast<uint32_t>(R), 3 * cin, kC);
      gx.c32 = w.dxcol;
      launch_gemm(gx, st);
      hipLaunchKernelGGL(col2im_kernel, blocks(R * cin), dim3(256), 0, st, w.dxcol, w.y[l - 1],
                         cin, R, w.dy, w.dy16);
      launch_check("col2im_kernel");
    }
  });
}

int gaffa_hip_scorer_adam(const gaffa_hip_scorer_net* net, float lr, float beta1, float beta2,
                          float eps, uint32_t step, gaffa_hip_stream stream) {
  return guarded([&] {
    check(net != nullptr && net->params != nullptr && net->params16 != nullptr &&
              net->grads != nullptr && net->adam_m != nullptr && net->adam_v != nullptr,
          "scorer Adam buffers must not be null");
    check(step >= 1, "Adam step must be >= 1");
    check(lr > 0.0f && beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f &&
              eps > 0.0f,
          "Adam hyperparameters out of range");
    const uint64_t n = make_layout().nparams;
    const float c1 = 1.0f - std::pow(beta1, static_cast<float>(step));
    const float c2 = 1.0f - std::pow(beta2, static_cast<float>(step));
    hipLaunchKernelGGL(adam_kernel, blocks(n), dim3(256), 0, as_stream(stream), net->params,
                       net->params16, net->grads, net->adam_m, net->adam_v, n, lr, beta1, beta2,
                       eps, c1, c2);
    launch_check("adam_kernel");
  });
}

int gaffa_hip_scorer_gather_patches(const gaffa_hip_scorer_patch_args* a,
                                    gaffa_hip_stream stream) {
  return guarded([&] {
    check(a != nullptr && a->periodogram != nullptr && a->series != nullptr &&
              a->row != nullptr && a->width != nullptr && a->out != nullptr,
          "scorer patch arguments must not be null");
    check(a->n > 0 && a->rows > 0 && a->widths > 0, "scorer patch sizes must be positive");
    check(a->series_stride >= a->rows * a->widths, "periodogram series stride too small");
    hipLaunchKernelGGL(gather_patch_kernel, dim3(a->n), dim3(kPatch), 0, as_stream(stream), *a);
    launch_check("gather_patch_kernel");
    hipLaunchKernelGGL(normalise_rows_kernel, dim3(a->n), dim3(256), 0, as_stream(stream),
                       a->out, kPatch, 0u);
    launch_check("normalise_rows_kernel");
  });
}

int gaffa_hip_scorer_normalise_rows(float* x, uint32_t rows, uint32_t length,
                                    gaffa_hip_stream stream) {
  return guarded([&] {
    check(x != nullptr && rows > 0 && length > 0, "normalise_rows needs a non-empty array");
    hipLaunchKernelGGL(normalise_rows_kernel, dim3(rows), dim3(256), 0, as_stream(stream), x,
                       length, 1u);
    launch_check("normalise_rows_kernel");
  });
}

int gaffa_hip_scorer_sync_params(const gaffa_hip_scorer_net* net, gaffa_hip_stream stream) {
  return guarded([&] {
    check(net != nullptr && net->params != nullptr && net->params16 != nullptr,
          "scorer parameters must not be null");
    const uint64_t n = make_layout().nparams;
    hipLaunchKernelGGL(sync_kernel, blocks(n), dim3(256), 0, as_stream(stream), net->params,
                       net->params16, n);
    launch_check("sync_kernel");
  });
}

}  // extern "C"
