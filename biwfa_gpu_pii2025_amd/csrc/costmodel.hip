// costmodel.hip -- learned pair-cost model for the BiWFA scheduler (gfx950).
//
// A 4 -> 64 -> 64 -> 1 ReLU MLP predicts a pair's alignment score from
//   x0 = log2(p+1)/16, x1 = log2(t+1)/16, x2 = log2(|p-t|+1)/16,
//   x3 = fraction of 16 pattern 12-mers (evenly spaced) found in the text near
//        their expected position,
// target y = log2(1 + score)/16.  Weights are stored in device memory as BF16
// (16-bit, round to nearest even of the FP32 master copy), widened to FP32 when a
// block stages them in LDS, and every product is accumulated in FP32.  The
// prediction only orders pairs and picks their starting workspace tier; it
// never changes an alignment result.
//
// Kernels: feature extraction (one wave per pair over the 2-bit packed batch),
// batched forward and per-block backward partial sums (the 64 x 64 products on
// FP32 MFMA tiles) reduced in a fixed order (deterministic gradients), and an
// Adam step.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "biwfa_mi355x.h"

namespace bwcm {

constexpr int kIn = 4, kH = 64;
// parameter vector layout (floats)
constexpr int kW1 = 0;                 // [kH][kIn]
constexpr int kB1 = kW1 + kH * kIn;    // [kH]
constexpr int kW2 = kB1 + kH;          // [kH][kH]  (row = output unit)
constexpr int kB2 = kW2 + kH * kH;     // [kH]
constexpr int kW3 = kB2 + kH;          // [kH]
constexpr int kB3 = kW3 + kH;          // [1]
constexpr int kParams = kB3 + 1;

// FP32 -> BF16 (round to nearest, ties to even) -> FP32
__host__ __device__ inline float bf16_round(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) == 0x7f800000u) return f;   // inf / nan unchanged
  const uint32_t lsb = (u >> 16) & 1u;
  u = (u + 0x7fffu + lsb) & 0xffff0000u;
  float r;
  std::memcpy(&r, &u, 4);
  return r;
}

// ---------------------------------------------------------------------------
// features

__device__ inline uint64_t window(const uint64_t* s, int pos) {   // bases [pos, pos+32)
  const int wi = pos >> 5;
  const int sh = (pos & 31) << 1;
  const uint64_t lo = s[wi], hi = s[wi + 1];
  return (lo >> sh) | ((hi << 1) << (63 - sh));
}

__global__ __launch_bounds__(256) void features_kernel(const uint64_t* words, const int64_t* p_woff,
                                                       const int32_t* p_len, const int64_t* t_woff,
                                                       const int32_t* t_len, int n, float* out) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= n) return;
  const int p = p_len[wave], t = t_len[wave];
  const uint64_t* P = words + p_woff[wave];
  const uint64_t* T = words + t_woff[wave];
  const int k = 12;
  const uint64_t kmask = (1ull << (2 * k)) - 1;
  int found = 0;
  if (p >= k && t >= k) {
    const int span = abs(p - t);
    const int W = 96 + span;
    for (int i = 0; i < 16; ++i) {
      const int pos = (int)(((long long)(2 * i + 1) * (p - k)) / 32);
      const uint64_t key = window(P, pos) & kmask;
      const int q = (int)(((long long)pos * (t - k)) / (p - k > 0 ? p - k : 1));
      const int lo = max(0, q - W), hi = min(t - k, q + W);
      bool hit = false;
      for (int j0 = lo; j0 <= hi && !hit; j0 += 64) {
        const int j = j0 + lane;
        const bool m = (j <= hi) && ((window(T, j) & kmask) == key);
        hit = __ballot(m) != 0ull;
      }
      found += hit ? 1 : 0;
    }
  }
  if (lane == 0) {
    float* o = out + (size_t)wave * kIn;
    o[0] = log2f((float)p + 1.0f) / 16.0f;
    o[1] = log2f((float)t + 1.0f) / 16.0f;
    o[2] = log2f((float)abs(p - t) + 1.0f) / 16.0f;
    o[3] = (p >= k && t >= k) ? (float)found / 16.0f : 0.0f;
  }
}

// ---------------------------------------------------------------------------
// MLP

// The upper half of an FP32 value that bf16_round left with a zero lower half.
__host__ __device__ inline uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (uint16_t)(u >> 16);
}
// A BF16 weight widened to FP32 (exact).
__host__ __device__ inline float bf16_value(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// Stores the FP32 master parameters rounded to BF16: the weights forward and
// backward use.
__global__ void to_bf16_kernel(const float* master, uint16_t* w, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) w[i] = bf16_bits(bf16_round(master[i]));
}

// The forward pass, with the 64 x 64 hidden layer on the matrix cores: each wave
// takes 32 samples at a time; layer 1 and the output layer run per lane, each
// unit's sum in input order, and layer 2 is two 32 x 32 FP32 MFMA tiles (32 steps
// of v_mfma_f32_32x32x2f32 each).  Lane l holds A[l % 32][l / 32] and
// B[l / 32][l % 32]; accumulator element v of lane l is row
// (v / 4) * 8 + (l / 32) * 4 + v % 4, column l % 32.  Products and sums stay FP32
// (the MFMA fixes layer 2's summation order).
constexpr int kMfmaWaves = 4;                 // waves per block
constexpr int kTileRows = 32;                 // samples per wave tile
constexpr int kPad = kH + 1;                  // padded row: conflict-free column reads
typedef float f32x16 __attribute__((ext_vector_type(16)));

__global__ __launch_bounds__(64 * kMfmaWaves) void forward_mfma_kernel(const uint16_t* w_bf16, const float* X,
                                                                       int n, float* Y) {
  __shared__ float w1[kH * kIn], b1[kH], b2[kH], w3[kH], b3;
  __shared__ float w2[kH * kPad];                         // [out][in], padded rows
  __shared__ float tile[kMfmaWaves][kTileRows * kPad];    // h1, then h2, of a wave's samples
  for (int i = threadIdx.x; i < kH * kIn; i += blockDim.x) w1[i] = bf16_value(w_bf16[kW1 + i]);
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    b1[i] = bf16_value(w_bf16[kB1 + i]);
    b2[i] = bf16_value(w_bf16[kB2 + i]);
    w3[i] = bf16_value(w_bf16[kW3 + i]);
  }
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) w2[(i / kH) * kPad + i % kH] = bf16_value(w_bf16[kW2 + i]);
  if (threadIdx.x == 0) b3 = bf16_value(w_bf16[kB3]);
  __syncthreads();
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int r = l & 31, half = l >> 5;
  float* t = tile[wv];
  for (int64_t s0 = ((int64_t)blockIdx.x * kMfmaWaves + wv) * kTileRows; s0 < n;
       s0 += (int64_t)gridDim.x * kMfmaWaves * kTileRows) {
    // layer 1: lane (r, half) computes units half*32 .. half*32+31 of sample s0+r
    float x[kIn] = {0.f, 0.f, 0.f, 0.f};
    if (s0 + r < n) {
#pragma unroll
      for (int c = 0; c < kIn; ++c) x[c] = X[(size_t)(s0 + r) * kIn + c];
    }
    for (int u = half * 32; u < half * 32 + 32; ++u) {
      float a = b1[u];
#pragma unroll
      for (int c = 0; c < kIn; ++c) a = fmaf(w1[u * kIn + c], x[c], a);
      t[r * kPad + u] = a > 0.f ? a : 0.f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // layer 2: D[sample][out] = sum_k h1[sample][k] * W2[out][k], two column blocks
    f32x16 acc0 = {}, acc1 = {};
    for (int kk = 0; kk < kH / 2; ++kk) {
      const int k = 2 * kk + half;
      const float a = t[r * kPad + k];
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w2[r * kPad + k], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w2[(32 + r) * kPad + k], acc1, 0, 0, 0);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // bias + ReLU into the tile (h1 is no longer read)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int row = (v / 4) * 8 + half * 4 + (v % 4);
      const float z0 = acc0[v] + b2[r], z1 = acc1[v] + b2[32 + r];
      t[row * kPad + r] = z0 > 0.f ? z0 : 0.f;
      t[row * kPad + 32 + r] = z1 > 0.f ? z1 : 0.f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // output layer, one sample per lane of the first half
    if (half == 0 && s0 + r < n) {
      float o = b3;
      for (int i = 0; i < kH; ++i) o = fmaf(w3[i], t[r * kPad + i], o);
      Y[s0 + r] = o;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// Gradient of the mean squared error over samples [0,n) w.r.t. the (BF16)
// weights; block b writes its partial sum over samples [64b, 64b+64).  Four
// waves; the three 64 x 64 products run as 32 x 32 FP32 MFMA tiles (lane layout
// as in forward_mfma_kernel):
//   forward   h2 = relu(h1 W2^T + b2)       waves 0-1, 32 samples each
//   backward  d1 = (d2 W2) masked by h1 > 0  one 32 x 32 tile per wave
//   gradient  dW2 = d2^T h1                 one 32 x 32 tile per wave
// where d2 = dL/d(pre-activation of layer 2).  Layer 1, the output layer and the
// bias and small-weight sums run per lane in sample order.  Deterministic: the
// same inputs give the same bits.
constexpr int kBwdRows = 64;                  // samples per block
__global__ __launch_bounds__(256) void backward_kernel(const uint16_t* w_bf16, const float* X, const float* Yt,
                                                       int n, float inv_n, float* partial) {
  __shared__ float w1[kH * kIn], b1[kH], w3[kH], b2[kH], b3;
  __shared__ float w2[kH * kPad];                 // [out][in], padded rows
  __shared__ float sx[kBwdRows][kIn];
  __shared__ float sh1[kBwdRows * kPad];          // layer-1 activations
  __shared__ float sh2[kBwdRows * kPad];          // layer-2 activations
  __shared__ float sd2[kBwdRows * kPad];          // dL/d(pre-activation of layer 2)
  __shared__ float sd1[kBwdRows * kPad];          // dL/d(pre-activation of layer 1)
  __shared__ float sg3[kBwdRows];                 // dL/dy
  for (int i = threadIdx.x; i < kH * kIn; i += blockDim.x) w1[i] = bf16_value(w_bf16[kW1 + i]);
  for (int i = threadIdx.x; i < kH; i += blockDim.x) {
    b1[i] = bf16_value(w_bf16[kB1 + i]);
    b2[i] = bf16_value(w_bf16[kB2 + i]);
    w3[i] = bf16_value(w_bf16[kW3 + i]);
  }
  for (int i = threadIdx.x; i < kH * kH; i += blockDim.x) w2[(i / kH) * kPad + i % kH] = bf16_value(w_bf16[kW2 + i]);
  if (threadIdx.x == 0) b3 = bf16_value(w_bf16[kB3]);
  __syncthreads();
  const int base = blockIdx.x * kBwdRows;
  const int tid = threadIdx.x, wv = tid >> 6, l = tid & 63, r = l & 31, half = l >> 5;
  if (wv < 2) {   // forward of samples base + 32 wv + (0 .. 31)
    const int s = 32 * wv + r;
    float x[kIn] = {0.f, 0.f, 0.f, 0.f};
    if (base + s < n) {
#pragma unroll
      for (int c = 0; c < kIn; ++c) x[c] = X[(size_t)(base + s) * kIn + c];
    }
    if (half == 0) {
#pragma unroll
      for (int c = 0; c < kIn; ++c) sx[s][c] = x[c];
    }
    for (int u = half * 32; u < half * 32 + 32; ++u) {
      float a = b1[u];
#pragma unroll
      for (int c = 0; c < kIn; ++c) a = fmaf(w1[u * kIn + c], x[c], a);
      sh1[s * kPad + u] = a > 0.f ? a : 0.f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    f32x16 acc0 = {}, acc1 = {};
    for (int kk = 0; kk < kH / 2; ++kk) {
      const int k = 2 * kk + half;
      const float a = sh1[(32 * wv + r) * kPad + k];
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w2[r * kPad + k], acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w2[(32 + r) * kPad + k], acc1, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int row = 32 * wv + (v / 4) * 8 + half * 4 + (v % 4);
      const float z0 = acc0[v] + b2[r], z1 = acc1[v] + b2[32 + r];
      sh2[row * kPad + r] = z0 > 0.f ? z0 : 0.f;
      sh2[row * kPad + 32 + r] = z1 > 0.f ? z1 : 0.f;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (half == 0) {   // output layer and dL/dy of sample s
      float o = b3;
      for (int i = 0; i < kH; ++i) o = fmaf(w3[i], sh2[s * kPad + i], o);
      sg3[s] = (base + s < n) ? 2.0f * (o - Yt[base + s]) * inv_n : 0.0f;
    }
  }
  __syncthreads();
  for (int q = tid; q < kBwdRows * kH; q += blockDim.x) {
    const int s = q / kH, j = q % kH;
    sd2[s * kPad + j] = (sh2[s * kPad + j] > 0.f) ? sg3[s] * w3[j] : 0.f;
  }
  __syncthreads();
  const int rb = wv & 1, cb = wv >> 1;   // this wave's 32 x 32 tile: row block, column block
  {  // d1[s][i] = sum_j d2[s][j] W2[j][i], masked by h1[s][i] > 0
    f32x16 acc = {};
    for (int kk = 0; kk < kH / 2; ++kk) {
      const int k = 2 * kk + half;
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sd2[(32 * rb + r) * kPad + k], w2[k * kPad + 32 * cb + r], acc, 0,
                                                 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int sr = 32 * rb + (v / 4) * 8 + half * 4 + (v % 4), i = 32 * cb + r;
      sd1[sr * kPad + i] = (sh1[sr * kPad + i] > 0.f) ? acc[v] : 0.f;
    }
  }
  float* out = partial + (size_t)blockIdx.x * kParams;
  {  // dW2[j][i] = sum_s d2[s][j] h1[s][i]
    f32x16 acc = {};
    for (int kk = 0; kk < kBwdRows / 2; ++kk) {
      const int k = 2 * kk + half;
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sd2[k * kPad + 32 * rb + r], sh1[k * kPad + 32 * cb + r], acc, 0,
                                                 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      const int j = 32 * rb + (v / 4) * 8 + half * 4 + (v % 4), i = 32 * cb + r;
      out[kW2 + j * kH + i] = acc[v];
    }
  }
  __syncthreads();
  // the other parameters' partial sums over the block's samples, in sample order
  for (int q = tid; q < kParams; q += blockDim.x) {
    if (q >= kW2 && q < kB2) continue;
    float acc = 0.f;
    if (q < kB1) {
      const int j = (q - kW1) / kIn, i = (q - kW1) % kIn;
      for (int s = 0; s < kBwdRows; ++s) acc = fmaf(sd1[s * kPad + j], sx[s][i], acc);
    } else if (q < kW2) {
      const int j = q - kB1;
      for (int s = 0; s < kBwdRows; ++s) acc += sd1[s * kPad + j];
    } else if (q < kW3) {
      const int j = q - kB2;
      for (int s = 0; s < kBwdRows; ++s) acc += sd2[s * kPad + j];
    } else if (q < kB3) {
      const int i = q - kW3;
      for (int s = 0; s < kBwdRows; ++s) acc = fmaf(sg3[s], sh2[s * kPad + i], acc);
    } else {
      for (int s = 0; s < kBwdRows; ++s) acc += sg3[s];
    }
    out[q] = acc;
  }
}

// grad[q] = sum over blocks of partial[b][q] (fixed order: deterministic)
__global__ void reduce_kernel(const float* partial, int blocks, float* grad) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kParams) return;
  float acc = 0.f;
  for (int b = 0; b < blocks; ++b) acc += partial[(size_t)b * kParams + q];
  grad[q] = acc;
}

__global__ void adam_kernel(float* master, const float* grad, float* m, float* v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kParams) return;
  const float g = grad[q];
  const float mq = b1 * m[q] + (1.f - b1) * g;
  const float vq = b2 * v[q] + (1.f - b2) * g * g;
  m[q] = mq;
  v[q] = vq;
  master[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

}  // namespace bwcm

using namespace bwcm;

struct bw_mlp {
  int dev = 0;
  hipStream_t stream = nullptr;
  float* d_master = nullptr;
  uint16_t* d_w = nullptr;    // BF16 weights (16-bit) used by forward/backward
  float* d_m = nullptr;
  float* d_v = nullptr;
  float* d_grad = nullptr;
  int64_t steps = 0;
};

namespace {
thread_local std::string g_err;

int fail(const char* what, hipError_t e) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return BW_STATUS_HIP_ERROR;
}

#define CM_TRY(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return fail(#expr, _e);      \
  } while (0)

template <class T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t n) { return hipMalloc(reinterpret_cast<void**>(&p), sizeof(T) * (n ? n : 1)); }
};

// Blocks of the forward kernel: one 32-sample tile per wave, at most 4096 blocks
// (the kernel strides over the rest).
unsigned forward_blocks(int64_t n) {
  const int64_t tiles = (n + kTileRows - 1) / kTileRows;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((tiles + kMfmaWaves - 1) / kMfmaWaves, 4096));
}

int refresh_bf16(bw_mlp* m) {
  hipLaunchKernelGGL(to_bf16_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, m->d_master,
                     m->d_w, kParams);
  CM_TRY(hipGetLastError());
  return 0;
}

// He-style initialisation from a seeded SplitMix64 stream (deterministic).
void init_params(std::vector<float>& w, uint64_t seed) {
  uint64_t x = seed * 0x9E3779B97F4A7C15ull + 12345;
  auto next = [&]() {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
  };
  auto uni = [&](double a) { return (float)((2.0 * next() - 1.0) * a); };
  w.assign(kParams, 0.f);
  for (int i = 0; i < kH * kIn; ++i) w[kW1 + i] = uni(std::sqrt(6.0 / kIn));
  for (int i = 0; i < kH; ++i) w[kB1 + i] = 0.05f;
  for (int i = 0; i < kH * kH; ++i) w[kW2 + i] = uni(std::sqrt(6.0 / kH));
  for (int i = 0; i < kH; ++i) w[kB2 + i] = 0.05f;
  for (int i = 0; i < kH; ++i) w[kW3 + i] = uni(std::sqrt(3.0 / kH));
  w[kB3] = 0.f;
}
}  // namespace

extern "C" {

int bw_mlp_num_params(void) { return kParams; }

bw_mlp_t* bw_mlp_new(int device, uint64_t seed) {
  bw_mlp* m = new bw_mlp();
  m->dev = device;
  std::vector<float> w;
  init_params(w, seed);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_master), 4 * kParams) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_w), sizeof(uint16_t) * kParams) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_m), 4 * kParams) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_v), 4 * kParams) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&m->d_grad), 4 * kParams) != hipSuccess ||
      hipMemcpy(m->d_master, w.data(), 4 * kParams, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(m->d_m, 0, 4 * kParams) != hipSuccess || hipMemset(m->d_v, 0, 4 * kParams) != hipSuccess ||
      refresh_bf16(m) != 0 || hipStreamSynchronize(m->stream) != hipSuccess) {
    g_err = "bw_mlp_new: device setup failed";
    bw_mlp_free(m);
    return nullptr;
  }
  return m;
}

void bw_mlp_free(bw_mlp_t* m) {
  if (!m) return;
  (void)hipSetDevice(m->dev);
  for (float* p : {m->d_master, m->d_m, m->d_v, m->d_grad})
    if (p) (void)hipFree(p);
  if (m->d_w) (void)hipFree(m->d_w);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

const char* bw_mlp_last_error(void) { return g_err.c_str(); }

int bw_mlp_get_params(bw_mlp_t* m, float* master, float* bf16_weights) {
  if (!m) return BW_STATUS_INVALID_ARG;
  CM_TRY(hipSetDevice(m->dev));
  if (master) CM_TRY(hipMemcpy(master, m->d_master, 4 * kParams, hipMemcpyDeviceToHost));
  if (bf16_weights) {
    std::vector<uint16_t> b(kParams);
    CM_TRY(hipMemcpy(b.data(), m->d_w, sizeof(uint16_t) * kParams, hipMemcpyDeviceToHost));
    for (int i = 0; i < kParams; ++i) bf16_weights[i] = bf16_value(b[i]);
  }
  return 0;
}

int bw_mlp_set_params(bw_mlp_t* m, const float* master) {
  if (!m || !master) return BW_STATUS_INVALID_ARG;
  CM_TRY(hipSetDevice(m->dev));
  CM_TRY(hipMemcpy(m->d_master, master, 4 * kParams, hipMemcpyHostToDevice));
  const int st = refresh_bf16(m);
  if (st) return st;
  CM_TRY(hipStreamSynchronize(m->stream));
  return 0;
}

// Predictions for n samples (host X[n][4] -> host y[n]).
int bw_mlp_predict(bw_mlp_t* m, const float* X, int64_t n, float* y) {
  if (!m || n < 0 || (n > 0 && (!X || !y))) return BW_STATUS_INVALID_ARG;
  if (n == 0) return 0;
  CM_TRY(hipSetDevice(m->dev));
  DevBuf<float> dX, dY;
  CM_TRY(dX.alloc(n * kIn));
  CM_TRY(dY.alloc(n));
  CM_TRY(hipMemcpyAsync(dX.p, X, 4 * n * kIn, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(forward_mfma_kernel, dim3(forward_blocks(n)), dim3(64 * kMfmaWaves), 0, m->stream, m->d_w,
                     dX.p, (int)n, dY.p);
  CM_TRY(hipGetLastError());
  CM_TRY(hipMemcpyAsync(y, dY.p, 4 * n, hipMemcpyDeviceToHost, m->stream));
  CM_TRY(hipStreamSynchronize(m->stream));
  return 0;
}

// Gradient of the mean squared error over all n samples (host), for checking.
int bw_mlp_gradient(bw_mlp_t* m, const float* X, const float* y, int64_t n, float* grad) {
  if (!m || n <= 0 || !X || !y || !grad) return BW_STATUS_INVALID_ARG;
  CM_TRY(hipSetDevice(m->dev));
  const int blocks = (int)((n + kBwdRows - 1) / kBwdRows);
  DevBuf<float> dX, dY, dP;
  CM_TRY(dX.alloc(n * kIn));
  CM_TRY(dY.alloc(n));
  CM_TRY(dP.alloc((size_t)blocks * kParams));
  CM_TRY(hipMemcpyAsync(dX.p, X, 4 * n * kIn, hipMemcpyHostToDevice, m->stream));
  CM_TRY(hipMemcpyAsync(dY.p, y, 4 * n, hipMemcpyHostToDevice, m->stream));
  hipLaunchKernelGGL(backward_kernel, dim3(blocks), dim3(256), 0, m->stream, m->d_w, dX.p, dY.p, (int)n,
                     1.0f / (float)n, dP.p);
  hipLaunchKernelGGL(reduce_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, dP.p, blocks,
                     m->d_grad);
  CM_TRY(hipGetLastError());
  CM_TRY(hipMemcpyAsync(grad, m->d_grad, 4 * kParams, hipMemcpyDeviceToHost, m->stream));
  CM_TRY(hipStreamSynchronize(m->stream));
  return 0;
}

// Mini-batch Adam training on the device.  Samples are visited in a seeded
// permutation per epoch; *final_loss receives the full-set MSE after training.
int bw_mlp_train(bw_mlp_t* m, const float* X, const float* y, int64_t n, int epochs, int batch, float lr,
                 uint64_t seed, float* final_loss) {
  if (!m || n <= 0 || !X || !y || epochs < 0 || batch <= 0) return BW_STATUS_INVALID_ARG;
  CM_TRY(hipSetDevice(m->dev));
  DevBuf<float> dX, dY, dXb, dYb, dP, dPred;
  CM_TRY(dX.alloc(n * kIn));
  CM_TRY(dY.alloc(n));
  CM_TRY(dXb.alloc((size_t)batch * kIn));
  CM_TRY(dYb.alloc(batch));
  const int max_blocks = (batch + kBwdRows - 1) / kBwdRows;
  CM_TRY(dP.alloc((size_t)max_blocks * kParams));
  CM_TRY(dPred.alloc(n));
  // host-side shuffled copies per epoch (the permutation is cheap; the math is on the device)
  std::vector<float> hx((size_t)batch * kIn), hy(batch);
  std::vector<int64_t> perm(n);
  for (int64_t i = 0; i < n; ++i) perm[i] = i;
  uint64_t st = seed * 0x9E3779B97F4A7C15ull + 7;
  auto rnd = [&]() {
    st ^= st << 13; st ^= st >> 7; st ^= st << 17;
    return st;
  };
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int ep = 0; ep < epochs; ++ep) {
    for (int64_t i = n - 1; i > 0; --i) std::swap(perm[i], perm[rnd() % (uint64_t)(i + 1)]);
    for (int64_t b0 = 0; b0 < n; b0 += batch) {
      const int nb = (int)std::min<int64_t>(batch, n - b0);
      for (int j = 0; j < nb; ++j) {
        const int64_t s = perm[b0 + j];
        for (int q = 0; q < kIn; ++q) hx[(size_t)j * kIn + q] = X[s * kIn + q];
        hy[j] = y[s];
      }
      CM_TRY(hipMemcpyAsync(dXb.p, hx.data(), 4 * (size_t)nb * kIn, hipMemcpyHostToDevice, m->stream));
      CM_TRY(hipMemcpyAsync(dYb.p, hy.data(), 4 * (size_t)nb, hipMemcpyHostToDevice, m->stream));
      const int blocks = (nb + kBwdRows - 1) / kBwdRows;
      hipLaunchKernelGGL(backward_kernel, dim3(blocks), dim3(256), 0, m->stream, m->d_w, dXb.p, dYb.p, nb,
                         1.0f / (float)nb, dP.p);
      hipLaunchKernelGGL(reduce_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, dP.p, blocks,
                         m->d_grad);
      m->steps++;
      const float c1 = 1.f - std::pow(b1, (float)m->steps), c2 = 1.f - std::pow(b2, (float)m->steps);
      hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->stream, m->d_master,
                         m->d_grad, m->d_m, m->d_v, lr, b1, b2, eps, c1, c2);
      const int r = refresh_bf16(m);
      if (r) return r;
      CM_TRY(hipGetLastError());
      CM_TRY(hipStreamSynchronize(m->stream));   // the host buffers are reused next batch
    }
  }
  if (final_loss) {
    std::vector<float> pred(n);
    CM_TRY(hipMemcpyAsync(dX.p, X, 4 * n * kIn, hipMemcpyHostToDevice, m->stream));
    hipLaunchKernelGGL(forward_mfma_kernel, dim3(forward_blocks(n)), dim3(64 * kMfmaWaves), 0, m->stream, m->d_w,
                       dX.p, (int)n, dPred.p);
    CM_TRY(hipGetLastError());
    CM_TRY(hipMemcpyAsync(pred.data(), dPred.p, 4 * n, hipMemcpyDeviceToHost, m->stream));
    CM_TRY(hipStreamSynchronize(m->stream));
    double acc = 0.0;
    for (int64_t i = 0; i < n; ++i) acc += (double)(pred[i] - y[i]) * (pred[i] - y[i]);
    *final_loss = (float)(acc / (double)n);
  }
  return 0;
}

// Features of every pair of a packed host batch (X[n][4], host).
int bw_batch_features(int device, const bw_batch_t* b, float* X) {
  if (!b || b->num_pairs < 0 || (b->num_pairs > 0 && !X)) return BW_STATUS_INVALID_ARG;
  const int64_t n = b->num_pairs;
  if (n == 0) return 0;
  if (n > INT32_MAX || !b->words || !b->pattern_word_offset || !b->text_word_offset || !b->pattern_length ||
      !b->text_length) {
    g_err = "bw_batch_features: invalid batch";
    return BW_STATUS_INVALID_ARG;
  }
  // the same layout rules as bw_dbatch_upload: one readable word before every
  // sequence and two after it, so the kernel's windows stay inside `words`
  for (int64_t i = 0; i < n; ++i) {
    const int p = b->pattern_length[i], t = b->text_length[i];
    const int64_t pw = b->pattern_word_offset[i], tw = b->text_word_offset[i];
    if (p < 0 || t < 0 || pw < 1 || tw < 1 || pw + (p + 31) / 32 + 2 > b->num_words ||
        tw + (t + 31) / 32 + 2 > b->num_words) {
      g_err = "bw_batch_features: pair " + std::to_string(i) + " has a bad length or word offset";
      return BW_STATUS_INVALID_ARG;
    }
  }
  CM_TRY(hipSetDevice(device));
  DevBuf<uint64_t> dW;
  DevBuf<int64_t> dPo, dTo;
  DevBuf<int32_t> dPl, dTl;
  DevBuf<float> dX;
  CM_TRY(dW.alloc(b->num_words));
  CM_TRY(dPo.alloc(n)); CM_TRY(dTo.alloc(n)); CM_TRY(dPl.alloc(n)); CM_TRY(dTl.alloc(n));
  CM_TRY(dX.alloc(n * kIn));
  CM_TRY(hipMemcpy(dW.p, b->words, 8 * b->num_words, hipMemcpyHostToDevice));
  CM_TRY(hipMemcpy(dPo.p, b->pattern_word_offset, 8 * n, hipMemcpyHostToDevice));
  CM_TRY(hipMemcpy(dTo.p, b->text_word_offset, 8 * n, hipMemcpyHostToDevice));
  CM_TRY(hipMemcpy(dPl.p, b->pattern_length, 4 * n, hipMemcpyHostToDevice));
  CM_TRY(hipMemcpy(dTl.p, b->text_length, 4 * n, hipMemcpyHostToDevice));
  const int waves_per_block = 4;
  hipLaunchKernelGGL(features_kernel, dim3((unsigned)((n + waves_per_block - 1) / waves_per_block)),
                     dim3(64 * waves_per_block), 0, 0, dW.p, dPo.p, dPl.p, dTo.p, dTl.p, (int)n, dX.p);
  CM_TRY(hipGetLastError());
  CM_TRY(hipMemcpy(X, dX.p, 4 * n * kIn, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
