// net.hip -- BF16 MFMA kernels of the learned preconditioner network (net.h).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "device_common.h"
#include "net.h"

#define HG_CHECK(x)                                                                  \
  do {                                                                               \
    hipError_t e_ = (x);                                                             \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP: ") +            \
                                                   hipGetErrorString(e_) + " at " #x); \
  } while (0)

namespace hg {

namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));  // one lane's 16x16x32 operand
using bf = uint16_t;

__device__ __forceinline__ bool Done(const int* d) { return d && *d; }

// FP32 -> BF16, round to nearest even.
__device__ __forceinline__ bf F2B(float f) {
  unsigned u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return bf(0x7fc0);
  u += 0x7fffu + ((u >> 16) & 1u);
  return bf(u >> 16);
}
__device__ __forceinline__ float B2F(bf b) { return __uint_as_float(unsigned(b) << 16); }

__device__ __forceinline__ void Unpack8(const uint4 u, float f[8]) {
  const unsigned w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    f[2 * k] = __uint_as_float(w[k] << 16);
    f[2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
  }
}
__device__ __forceinline__ uint4 Pack8(const float f[8]) {
  unsigned w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) w[k] = unsigned(F2B(f[2 * k])) | (unsigned(F2B(f[2 * k + 1])) << 16);
  return uint4{w[0], w[1], w[2], w[3]};
}
__device__ __forceinline__ uint4 Ld8(const bf* p) { return *reinterpret_cast<const uint4*>(p); }
// 8 BF16 values (16 bytes, 16-byte aligned) as an MFMA operand
__device__ __forceinline__ b8 LdOp(const bf* p) { return __builtin_bit_cast(b8, Ld8(p)); }
__device__ __forceinline__ void St8(bf* p, uint4 v) { *reinterpret_cast<uint4*>(p) = v; }


// ---------------------------------------------------------------------------
// 3x3 'same' convolution, CIN, COUT multiples of 16, as an implicit GEMM on
// gfx950's MFMA 16x16x32 BF16 (FP32 accumulate).  A block of 4 waves
// computes a 64 x 4 pixel tile (one row per wave) for all output channels.
// The input tile with its one-pixel halo is staged once in LDS (16-byte
// global loads and LDS stores; pixels padded to CIN + 8 channels, a 16-byte
// multiple, so every operand read is one aligned 16-byte LDS load).  The
// GEMM's K runs over (tap, input channel), k = tap * CIN + ci, 32 at a time:
// per chunk a wave multiplies each weight block (M = 16 output channels,
// read once and used kSeg times) by kSeg input blocks from LDS (N = 16
// pixels each); with CIN = 16 a chunk spans two taps, and the chunk past
// the ninth tap is zero.  Operand lane layouts (checked on the device):
// A[i][k] in lane i + 16*(k/8), element k%8 (8 consecutive channels of one
// tap: 16 contiguous bytes of wt[tap][co][ci]); B[k][j] in lane j +
// 16*(k/8), element k%8; D[i][j] in lane j + 16*(i/4), element i%4, so each
// lane ends with 4 consecutive output channels of one pixel and stores them
// as one 8-byte word.  Used for the forward pass and, with tap-flipped
// weights, for dgrad.
constexpr int kSeg = 4;           // 16-pixel blocks per wave: 64-pixel rows
constexpr int kCTX = 16 * kSeg;   // tile width
constexpr int kCTY = 4;           // tile height (one row per wave)

template <int CIN, int COUT, bool RELU>
__global__ void __launch_bounds__(256)
KConvMfma(int W, int H, const bf* __restrict__ in, const bf* __restrict__ wt,
          bf* __restrict__ out, const int* done) {
  if (Done(done)) return;
  static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel counts");
  constexpr int CP = CIN + 8;  // padded channel stride in LDS (16-byte multiple)
  constexpr int HX = kCTX + 2, HY = kCTY + 2;
  __shared__ __attribute__((aligned(16))) bf sin_[HY * HX * CP];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int tiles_x = (W + kCTX - 1) / kCTX;
  const int bx = blockIdx.x % tiles_x, by = blockIdx.x / tiles_x;
  const int x0 = bx * kCTX, y0 = by * kCTY;
  // stage the input tile + halo: 8 channels (16 bytes) per load
  constexpr int G = CIN / 8;
  for (int q = threadIdx.x; q < HX * HY * G; q += 256) {
    const int g = q % G, pix = q / G;
    const int hx = pix % HX, hy = pix / HX;
    const int xx = x0 + hx - 1, yy = y0 + hy - 1;
    uint4 v = uint4{0u, 0u, 0u, 0u};
    if (xx >= 0 && xx < W && yy >= 0 && yy < H) v = Ld8(in + (size_t(yy) * W + xx) * CIN + g * 8);
    St8(sin_ + pix * CP + g * 8, v);
  }
  __syncthreads();
  const int ty = y0 + wave;
  if (ty >= H) return;
  const int pj = lane & 15, kq = lane >> 4;
  f4 acc[kSeg][COUT / 16];
#pragma unroll
  for (int s = 0; s < kSeg; ++s)
#pragma unroll
    for (int n = 0; n < COUT / 16; ++n) acc[s][n] = f4{0.f, 0.f, 0.f, 0.f};
  constexpr int NQ = (9 * CIN + 31) / 32;  // K chunks
  const b8 zero = __builtin_bit_cast(b8, uint4{0u, 0u, 0u, 0u});
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    // this lane's 8 values of the chunk: channels ci0..ci0+7 of tap t
    const int kk = 32 * q + 8 * kq;
    const int t = kk / CIN, ci0 = kk % CIN;
    const bool live = t < 9;
    const int hy = wave + t / 3;
    b8 bw[COUT / 16];
#pragma unroll
    for (int n = 0; n < COUT / 16; ++n)
      bw[n] = live ? LdOp(wt + size_t(t * COUT + n * 16 + pj) * CIN + ci0) : zero;
#pragma unroll
    for (int s = 0; s < kSeg; ++s) {
      const int hx = 16 * s + pj + t % 3;
      const b8 xin = live ? LdOp(sin_ + (hy * HX + hx) * CP + ci0) : zero;
#pragma unroll
      for (int n = 0; n < COUT / 16; ++n)
        acc[s][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[n], xin, acc[s][n], 0, 0, 0);
    }
  }
#pragma unroll
  for (int s = 0; s < kSeg; ++s) {
    const int px = x0 + 16 * s + pj;
    if (px >= W) continue;
#pragma unroll
    for (int n = 0; n < COUT / 16; ++n) {
      s4 o;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float r = acc[s][n][v];
        if (RELU) r = fmaxf(r, 0.f);
        o[v] = short(F2B(r));
      }
      *reinterpret_cast<s4*>(out + (size_t(ty) * W + px) * COUT + n * 16 + kq * 4) = o;
    }
  }
}

// Weight gradient dW[tap][ci][co] += sum_p in[p + off(tap)][ci] g[p][co]:
// per tap an implicit GEMM with M = 16 input channels, N = COUT, K = pixels,
// 32 pixels per MFMA 16x16x32 BF16.  Grid: x = tap * (CIN/16) + channel
// tile, y = row blocks of `rows` rows.  The block walks its rows in chunks of
// kWgCols pixels: it stages the chunk's input pixels (shifted by the tap's
// offset, zero outside the grid; 16 channels) and output gradients (COUT
// channels) in LDS channel-major, from 16-byte global loads, so that each
// lane's operand -- 8 consecutive pixels of one channel -- is one aligned
// 16-byte LDS load; wave w takes the chunk's pixels 32 w .. 32 w + 31.
constexpr int kWgCols = 128;                 // pixels per staged chunk (4 waves x 32)
constexpr int kWgStride = kWgCols + 8;       // LDS row stride (16-byte multiple)
template <int CIN, int COUT>
__global__ void __launch_bounds__(256)
KConvWgrad(int W, int H, int rows, const bf* __restrict__ in, const bf* __restrict__ g,
           float* __restrict__ dw) {
  static_assert(CIN % 16 == 0 && COUT % 16 == 0, "channel counts");
  __shared__ __attribute__((aligned(16))) bf sa[16 * kWgStride];
  __shared__ __attribute__((aligned(16))) bf sb[COUT * kWgStride];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int t = blockIdx.x / (CIN / 16), ct = blockIdx.x % (CIN / 16);
  const int dy = t / 3 - 1, dx = t % 3 - 1;
  const int kq = lane >> 4, pj = lane & 15;
  f4 acc[COUT / 16];
#pragma unroll
  for (int n = 0; n < COUT / 16; ++n) acc[n] = f4{0.f, 0.f, 0.f, 0.f};
  const int y0 = blockIdx.y * rows, y1 = min(H, y0 + rows);
  for (int y = y0; y < y1; ++y) {
    const int yy = y + dy;
    for (int x0 = 0; x0 < W; x0 += kWgCols) {
      // input: 16 channels = two 8-channel loads per pixel
      for (int q = threadIdx.x; q < kWgCols * 2; q += 256) {
        const int px = q >> 1, half = q & 1;
        const int x = x0 + px, xx = x + dx;
        uint4 v = uint4{0u, 0u, 0u, 0u};
        if (x < W && xx >= 0 && xx < W && yy >= 0 && yy < H)
          v = Ld8(in + (size_t(yy) * W + xx) * CIN + ct * 16 + half * 8);
        float f[8];
        Unpack8(v, f);
#pragma unroll
        for (int e = 0; e < 8; ++e) sa[(half * 8 + e) * kWgStride + px] = bf(__float_as_uint(f[e]) >> 16);
      }
      // output gradient: COUT channels, 8 per load
      for (int q = threadIdx.x; q < kWgCols * (COUT / 8); q += 256) {
        const int px = q / (COUT / 8), grp = q % (COUT / 8);
        const int x = x0 + px;
        uint4 v = uint4{0u, 0u, 0u, 0u};
        if (x < W) v = Ld8(g + (size_t(y) * W + x) * COUT + grp * 8);
        float f[8];
        Unpack8(v, f);
#pragma unroll
        for (int e = 0; e < 8; ++e) sb[(grp * 8 + e) * kWgStride + px] = bf(__float_as_uint(f[e]) >> 16);
      }
      __syncthreads();
      const int k0 = 32 * wave + 8 * kq;  // this lane's 8 pixels of the wave's 32
      const b8 a = LdOp(sa + pj * kWgStride + k0);
#pragma unroll
      for (int n = 0; n < COUT / 16; ++n)
        acc[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, LdOp(sb + (n * 16 + pj) * kWgStride + k0),
                                                         acc[n], 0, 0, 0);
      __syncthreads();  // the chunk's LDS is restaged next
    }
  }
  // reduce the four waves in LDS, then one atomic add per weight
  __shared__ float red[4][COUT / 16][4][64];
#pragma unroll
  for (int n = 0; n < COUT / 16; ++n)
#pragma unroll
    for (int v = 0; v < 4; ++v) red[wave][n][v][lane] = acc[n][v];
  __syncthreads();
  for (int q = threadIdx.x; q < (COUT / 16) * 4 * 64; q += 256) {
    const int n = q / 256, v = (q / 64) % 4, l = q % 64;
    const float s = red[0][n][v][l] + red[1][n][v][l] + red[2][n][v][l] + red[3][n][v][l];
    // D[i][j]: i = 4*(l/16) + v (input channel), j = l % 16 (output channel)
    const int i = 4 * (l >> 4) + v, j = l & 15;
    atomicAdd(&dw[(size_t(t) * CIN + ct * 16 + i) * COUT + n * 16 + j], s);
  }
}

// E0: 1 -> 16 channels, FP32 input, ReLU, BF16 output; w[tap][co].
__global__ void __launch_bounds__(256)
KConvE0(int W, int H, const float* __restrict__ x, const float* __restrict__ w,
        bf* __restrict__ out, bool relu, const int* done) {
  if (Done(done)) return;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= W * H) return;
  const int px = p % W, py = p / W;
  float acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = 0.f;
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int xx = px + t % 3 - 1, yy = py + t / 3 - 1;
    if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
    const float v = x[yy * W + xx];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] += v * w[t * 16 + c];
  }
  if (relu)
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] = fmaxf(acc[c], 0.f);
  St8(out + size_t(p) * 16, Pack8(acc));
  St8(out + size_t(p) * 16 + 8, Pack8(acc + 8));
}

// D0: 16 -> 1 channel, FP32 output y = base + conv(in); w[tap][ci].  y may
// alias base (each thread reads its base value before writing it).
__global__ void __launch_bounds__(256)
KConvD0(int W, int H, const bf* __restrict__ in, const float* __restrict__ w,
        const float* base, float* y, const int* done) {
  if (Done(done)) return;
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= W * H) return;
  const int px = p % W, py = p / W;
  float acc = 0.f;
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int xx = px + t % 3 - 1, yy = py + t / 3 - 1;
    if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
    const bf* q = in + (size_t(yy) * W + xx) * 16;
    float f[16];
    Unpack8(Ld8(q), f);
    Unpack8(Ld8(q + 8), f + 8);
#pragma unroll
    for (int c = 0; c < 16; ++c) acc += f[c] * w[t * 16 + c];
  }
  y[p] = base ? base[p] + acc : acc;
}

// D0 dgrad: gin[p][ci] = sum_t gy[p - off(t)] w[t][ci].
__global__ void __launch_bounds__(256)
KConvD0Dgrad(int W, int H, const float* __restrict__ gy, const float* __restrict__ w,
             bf* __restrict__ gin) {
  const int p = blockIdx.x * 256 + threadIdx.x;
  if (p >= W * H) return;
  const int px = p % W, py = p / W;
  float acc[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) acc[c] = 0.f;
#pragma unroll
  for (int t = 0; t < 9; ++t) {
    const int xx = px - (t % 3 - 1), yy = py - (t / 3 - 1);
    if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
    const float g = gy[yy * W + xx];
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[c] += g * w[t * 16 + c];
  }
  St8(gin + size_t(p) * 16, Pack8(acc));
  St8(gin + size_t(p) * 16 + 8, Pack8(acc + 8));
}

// Weight gradients of the 1-channel ends, reduced per block then atomics:
// E0: dw[t][co] += x[p + off(t)] g[p][co];  D0: dw[t][ci] += in[p + off(t)][ci] gy[p].
template <bool E0>
__global__ void __launch_bounds__(256)
KEndWgrad(int W, int H, const float* __restrict__ xs, const bf* __restrict__ gb,
          const bf* __restrict__ inb, const float* __restrict__ gys, float* __restrict__ dw) {
  __shared__ float red[256][17];
  float acc[9][16];
#pragma unroll
  for (int t = 0; t < 9; ++t)
#pragma unroll
    for (int c = 0; c < 16; ++c) acc[t][c] = 0.f;
  for (int p = blockIdx.x * 256 + threadIdx.x; p < W * H; p += gridDim.x * 256) {
    const int px = p % W, py = p / W;
#pragma unroll
    for (int t = 0; t < 9; ++t) {
      const int xx = px + t % 3 - 1, yy = py + t / 3 - 1;
      if (xx < 0 || xx >= W || yy < 0 || yy >= H) continue;
      if (E0) {
        const float v = xs[yy * W + xx];
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[t][c] += v * B2F(gb[size_t(p) * 16 + c]);
      } else {
        const float g = gys[p];
#pragma unroll
        for (int c = 0; c < 16; ++c) acc[t][c] += B2F(inb[(size_t(yy) * W + xx) * 16 + c]) * g;
      }
    }
  }
  for (int t = 0; t < 9; ++t) {
    for (int c = 0; c < 16; ++c) red[threadIdx.x][c] = acc[t][c];
    __syncthreads();
    if (threadIdx.x < 16) {
      float s = 0.f;
      for (int q = 0; q < 256; ++q) s += red[q][threadIdx.x];
      atomicAdd(&dw[t * 16 + threadIdx.x], s);
    }
    __syncthreads();
  }
}

// ---- elementwise ops on BF16 NHWC tensors ----------------------------------
// Channel counts are multiples of 16: every thread moves groups of 8 channels
// (16 bytes).  Sums are formed in FP32 in a fixed order and rounded once.
__global__ void KPool(int W, int H, int C, const bf* __restrict__ in, bf* __restrict__ out,
                      const int* done) {
  if (Done(done)) return;
  const int Wc = W / 2, Hc = H / 2, G = C / 8;
  const long n = long(Wc) * Hc * G;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += gridDim.x * 256L) {
    const int c = int(q % G) * 8;
    const long p = q / G;
    const int X = int(p % Wc), Y = int(p / Wc);
    const bf* b = in + (size_t(2 * Y) * W + 2 * X) * C + c;
    float f0[8], f1[8], f2[8], f3[8], o[8];
    Unpack8(Ld8(b), f0);
    Unpack8(Ld8(b + C), f1);
    Unpack8(Ld8(b + size_t(W) * C), f2);
    Unpack8(Ld8(b + size_t(W) * C + C), f3);
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = 0.25f * (f0[k] + f1[k] + f2[k] + f3[k]);
    St8(out + p * C + c, Pack8(o));
  }
}
// out = up(d) + e (fine grid W x H, C channels)
__global__ void KUpAdd(int W, int H, int C, const bf* __restrict__ d, const bf* __restrict__ e,
                       bf* __restrict__ out, const int* done) {
  if (Done(done)) return;
  const int Wc = W / 2, G = C / 8;
  const long n = long(W) * H * G;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += gridDim.x * 256L) {
    const int c = int(q % G) * 8;
    const long p = q / G;
    const int x = int(p % W), y = int(p / W);
    float fd[8], fe[8], o[8];
    Unpack8(Ld8(d + (size_t(y / 2) * Wc + x / 2) * C + c), fd);
    Unpack8(Ld8(e + p * C + c), fe);
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = fd[k] + fe[k];
    St8(out + p * C + c, Pack8(o));
  }
}
// g_coarse = sum of the four children of g_fine
__global__ void KUpBwd(int Wc, int Hc, int C, const bf* __restrict__ gf, bf* __restrict__ gc) {
  const int W = 2 * Wc, G = C / 8;
  const long n = long(Wc) * Hc * G;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += gridDim.x * 256L) {
    const int c = int(q % G) * 8;
    const long p = q / G;
    const int X = int(p % Wc), Y = int(p / Wc);
    const bf* b = gf + (size_t(2 * Y) * W + 2 * X) * C + c;
    float f0[8], f1[8], f2[8], f3[8], o[8];
    Unpack8(Ld8(b), f0);
    Unpack8(Ld8(b + C), f1);
    Unpack8(Ld8(b + size_t(W) * C), f2);
    Unpack8(Ld8(b + size_t(W) * C + C), f3);
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = f0[k] + f1[k] + f2[k] + f3[k];
    St8(gc + p * C + c, Pack8(o));
  }
}
// g_fine (+)= 0.25 * g_coarse at each child
__global__ void KPoolBwdAdd(int W, int H, int C, const bf* __restrict__ gc, bf* __restrict__ gf) {
  const int Wc = W / 2, G = C / 8;
  const long n = long(W) * H * G;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += gridDim.x * 256L) {
    const int c = int(q % G) * 8;
    const long p = q / G;
    const int x = int(p % W), y = int(p / W);
    float ff[8], fc[8], o[8];
    Unpack8(Ld8(gf + p * C + c), ff);
    Unpack8(Ld8(gc + (size_t(y / 2) * Wc + x / 2) * C + c), fc);
#pragma unroll
    for (int k = 0; k < 8; ++k) o[k] = ff[k] + 0.25f * fc[k];
    St8(gf + p * C + c, Pack8(o));
  }
}
// out = g * (act > 0); n is a multiple of 8
__global__ void KMask(long n, const bf* __restrict__ g, const bf* __restrict__ act,
                      bf* __restrict__ out) {
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n / 8; q += gridDim.x * 256L) {
    const uint4 gv = Ld8(g + q * 8), av = Ld8(act + q * 8);
    const unsigned gw[4] = {gv.x, gv.y, gv.z, gv.w}, aw[4] = {av.x, av.y, av.z, av.w};
    unsigned ow[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const unsigned lo = __uint_as_float(aw[k] << 16) > 0.f ? (gw[k] & 0xffffu) : 0u;
      const unsigned hi = __uint_as_float(aw[k] & 0xffff0000u) > 0.f ? (gw[k] & 0xffff0000u) : 0u;
      ow[k] = lo | hi;
    }
    St8(out + q * 8, uint4{ow[0], ow[1], ow[2], ow[3]});
  }
}

// loss = mean((y - t)^2) (partials), gy = 2 (y - t) / N
__global__ void __launch_bounds__(256)
KLoss(long n, const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ gy,
      double* __restrict__ part) {
  double s = 0.;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += gridDim.x * 256L) {
    const float d = y[q] - t[q];
    gy[q] = 2.f * d / float(n);
    s += double(d) * double(d);
  }
  s = WaveSum(s);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// Energy-norm loss: d = o*y - t (o: optional per-cell output scale), q = A d
// (face-form operator, FP32 coefficients kx[(nx+1)*ny], ky[nx*(ny+1)]);
// loss = d.q / n (partials), gy = o * 2 q / n.
__global__ void KDiff(long n, float* __restrict__ y, const float* __restrict__ t,
                      const float* __restrict__ o) {
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += gridDim.x * 256L)
    y[q] = (o ? o[q] * y[q] : y[q]) - t[q];
}

__global__ void __launch_bounds__(256)
KEnergyLoss(int nx, int ny, const float* __restrict__ kx, const float* __restrict__ ky,
            const float* __restrict__ d, const float* __restrict__ o, float* __restrict__ gy,
            double* __restrict__ part) {
  const long n = long(nx) * ny;
  double s = 0.;
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) {
    const int i = int(c % nx), j = int(c / nx);
    const float kw = kx[i + long(nx + 1) * j], ke = kx[i + 1 + long(nx + 1) * j];
    const float ks = ky[i + long(nx) * j], kn = ky[i + long(nx) * (j + 1)];
    float q = (((kw + ke) + ks) + kn) * d[c];
    if (i > 0) q -= kw * d[c - 1];
    if (i < nx - 1) q -= ke * d[c + 1];
    if (j > 0) q -= ks * d[c - nx];
    if (j < ny - 1) q -= kn * d[c + nx];
    gy[c] = (o ? o[c] : 1.f) * (2.f * q / float(n));
    s += double(d[c]) * double(q);
  }
  s = WaveSum(s);
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void KAdam(int n, float* __restrict__ w, const float* __restrict__ g,
                      float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                      float c1, float c2) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const float gg = g[q];
  const float mm = b1 * m[q] + (1.f - b1) * gg;
  const float vv = b2 * v[q] + (1.f - b2) * gg * gg;
  m[q] = mm;
  v[q] = vv;
  w[q] -= lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
}

// Scale-free energy loss: u = o*y, c = u.(A t), Y = u.(A u), T = t.(A t)
// (per-block partials part[3b..3b+2]; qu = A u and qt = A t kept).
__global__ void __launch_bounds__(256)
KCosDots(int nx, int ny, const float* __restrict__ kx, const float* __restrict__ ky,
         const float* __restrict__ y, const float* __restrict__ o, const float* __restrict__ t,
         float* __restrict__ qu, float* __restrict__ qt, double* __restrict__ part) {
  const long n = long(nx) * ny;
  double sc = 0., sy = 0., st = 0.;
  auto U = [&](long c) { return o ? o[c] * y[c] : y[c]; };
  for (long c = blockIdx.x * 256L + threadIdx.x; c < n; c += gridDim.x * 256L) {
    const int i = int(c % nx), j = int(c / nx);
    const float kw = kx[i + long(nx + 1) * j], ke = kx[i + 1 + long(nx + 1) * j];
    const float ks = ky[i + long(nx) * j], kn = ky[i + long(nx) * (j + 1)];
    const float dg = ((kw + ke) + ks) + kn;
    const float uc = U(c), tc = t[c];
    float au = dg * uc, at = dg * tc;
    if (i > 0) au -= kw * U(c - 1), at -= kw * t[c - 1];
    if (i < nx - 1) au -= ke * U(c + 1), at -= ke * t[c + 1];
    if (j > 0) au -= ks * U(c - nx), at -= ks * t[c - nx];
    if (j < ny - 1) au -= kn * U(c + nx), at -= kn * t[c + nx];
    qu[c] = au;
    qt[c] = at;
    sc += double(uc) * at;
    sy += double(uc) * au;
    st += double(tc) * at;
  }
  sc = WaveSum(sc);
  sy = WaveSum(sy);
  st = WaveSum(st);
  __shared__ double red[3][4];
  if ((threadIdx.x & 63) == 0) {
    red[0][threadIdx.x >> 6] = sc;
    red[1][threadIdx.x >> 6] = sy;
    red[2][threadIdx.x >> 6] = st;
  }
  __syncthreads();
  if (threadIdx.x < 3)
    part[3 * blockIdx.x + threadIdx.x] =
        red[threadIdx.x][0] + red[threadIdx.x][1] + red[threadIdx.x][2] + red[threadIdx.x][3];
}

// Gradient of L = 1 - c^2 / (Y T) (one minus the squared energy-cosine of u
// and t, i.e. the relative energy error of the best multiple alpha = c/Y of
// u): dL/du = (2 c / (Y T)) (alpha A u - A t); gy = o * dL/du.  Every block
// folds the partials in the same order; block 0 stores (L, alpha) in stat.
__global__ void __launch_bounds__(256)
KCosGrad(long n, int nb, const double* __restrict__ part, const float* __restrict__ o,
         const float* __restrict__ qu, const float* __restrict__ qt, float* __restrict__ gy,
         double* __restrict__ stat) {
  __shared__ double red[3][4];
  double a[3] = {0., 0., 0.};
  for (int b = threadIdx.x; b < nb; b += 256)
    for (int k = 0; k < 3; ++k) a[k] += part[3 * b + k];
  for (int k = 0; k < 3; ++k) {
    const double v = WaveSum(a[k]);
    if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  const double c = red[0][0] + red[0][1] + red[0][2] + red[0][3];
  const double Y = red[1][0] + red[1][1] + red[1][2] + red[1][3];
  const double T = red[2][0] + red[2][1] + red[2][2] + red[2][3];
  const bool ok = Y > 0. && T > 0.;
  const double alpha = ok ? c / Y : 0.;
  const float f = ok ? float(2. * c / (Y * T)) : 0.f, al = float(alpha);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    stat[0] = ok ? 1. - c * c / (Y * T) : 1.;
    stat[1] = alpha;
  }
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += gridDim.x * 256L)
    gy[q] = (o ? o[q] : 1.f) * (f * (al * qu[q] - qt[q]));
}

// w[0..n) *= a
__global__ void KScaleW(int n, float* __restrict__ w, float a) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q < n) w[q] *= a;
}

// Tied decoder of the first layer: wd0[t][c] = wE0[8 - t][c].
__global__ void KFlipE0(const float* __restrict__ w, float* __restrict__ wd0) {
  const int q = threadIdx.x;
  if (q < 144) wd0[q] = w[(8 - q / 16) * 16 + q % 16];
}

// Gradient of a tied decoder layer (layout [t][a][b], a < cout, b < cin of
// the encoder layer it mirrors) added to the encoder layer's gradient
// gE[8 - t][b][a].
__global__ void KFoldTied(int cin, int cout, const float* __restrict__ gd, float* __restrict__ ge) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  if (q >= 9 * cin * cout) return;
  const int t = q / (cin * cout), r = q % (cin * cout), a = r / cin, b = r % cin;
  ge[(size_t(8 - t) * cin + b) * cout + a] += gd[q];
}

// Master W[tap][ci][co] -> forward wf[tap][co][ci] and dgrad
// wb[tap][ci][co] = W[8 - tap][ci][co] (BF16).
__global__ void KConvertW(int cin, int cout, const float* __restrict__ w, bf* __restrict__ wf,
                          bf* __restrict__ wb) {
  const int q = blockIdx.x * 256 + threadIdx.x;
  const int n = 9 * cin * cout;
  if (q >= n) return;
  const int t = q / (cin * cout), r = q % (cin * cout), ci = r / cout, co = r % cout;
  const bf v = F2B(w[q]);
  wf[(size_t(t) * cout + co) * cin + ci] = v;
  wb[(size_t(8 - t) * cin + ci) * cout + co] = v;
}

inline unsigned Blk(long n) { return unsigned(std::min<long>((n + 255) / 256, 65535)); }

// Dispatch of the MFMA conv for the (cin, cout) pairs of the network.
void ConvFwd(int cin, int cout, bool relu, int W, int H, const bf* in, const bf* wt, bf* out,
             hipStream_t st, const int* done) {
  const unsigned grid = unsigned(((W + kCTX - 1) / kCTX) * ((H + kCTY - 1) / kCTY));
#define HG_CONV(CI, CO)                                                                  \
  if (cin == CI && cout == CO) {                                                         \
    if (relu)                                                                            \
      KConvMfma<CI, CO, true><<<grid, 256, 0, st>>>(W, H, in, wt, out, done);            \
    else                                                                                 \
      KConvMfma<CI, CO, false><<<grid, 256, 0, st>>>(W, H, in, wt, out, done);           \
    return;                                                                              \
  }
  HG_CONV(16, 32) HG_CONV(32, 64) HG_CONV(64, 64) HG_CONV(64, 32) HG_CONV(32, 16)
  HG_CONV(16, 16) HG_CONV(32, 32) HG_CONV(16, 64) HG_CONV(64, 16)
#undef HG_CONV
  throw std::runtime_error("no MFMA conv for this channel pair");
}

void ConvWgrad(int cin, int cout, int W, int H, const bf* in, const bf* g, float* dw,
               hipStream_t st) {
  const int rows = std::max(1, std::min(H, 8192 / std::max(W, 1)));
  dim3 grid(unsigned(9 * (cin / 16)), unsigned((H + rows - 1) / rows));
#define HG_WG(CI, CO)                                                                     \
  if (cin == CI && cout == CO) {                                                          \
    KConvWgrad<CI, CO><<<grid, 256, 0, st>>>(W, H, rows, in, g, dw);                      \
    return;                                                                               \
  }
  HG_WG(16, 32) HG_WG(32, 64) HG_WG(64, 64) HG_WG(64, 32) HG_WG(32, 16)
  HG_WG(16, 16) HG_WG(32, 32) HG_WG(16, 64) HG_WG(64, 16)
#undef HG_WG
  throw std::runtime_error("no MFMA wgrad for this channel pair");
}

}  // namespace

// ---------------------------------------------------------------------------

UNet::UNet(int nx, int ny) : nx_(nx), ny_(ny) {
  if (!Supported(nx, ny)) throw std::runtime_error("UNet: grid dimensions must be multiples of 8");
  for (int l = 0; l < kLevels; ++l) {
    lw_[l] = nx >> l;
    lh_[l] = ny >> l;
    ch_[l] = std::min(16 << l, 64);
  }
  ch_[kLevels - 1] = ch_[kLevels - 2];
  int off = 0, k = 0;
  layer_[k++] = {off, 1, ch_[0], 0};
  off += 9 * 1 * ch_[0];
  for (int l = 1; l < kLevels; ++l) {
    layer_[k++] = {off, ch_[l - 1], ch_[l], l};
    off += 9 * ch_[l - 1] * ch_[l];
  }
  for (int l = kLevels - 1; l >= 1; --l) {
    layer_[k++] = {off, ch_[l], ch_[l - 1], l};
    off += 9 * ch_[l] * ch_[l - 1];
  }
  layer_[k++] = {off, ch_[0], 1, 0};
  off += 9 * ch_[0];
  nw_ = off;
  auto A = [](void** p, size_t bytes) {
    HG_CHECK(hipMalloc(p, std::max<size_t>(bytes, 16)));
    HG_CHECK(hipMemset(*p, 0, std::max<size_t>(bytes, 16)));
  };
  A((void**)&w_, size_t(nw_) * 4);
  A((void**)&g_, size_t(nw_) * 4);
  A((void**)&m_, size_t(nw_) * 4);
  A((void**)&v_, size_t(nw_) * 4);
  A((void**)&wf_, size_t(nw_) * 2);
  A((void**)&wb_, size_t(nw_) * 2);
  A((void**)&wd0_, 9 * 16 * 4);
  A((void**)&gtie_, size_t(9) * 64 * 64 * 4);
  for (int l = 0; l < kLevels; ++l) {
    const size_t px = size_t(lw_[l]) * lh_[l];
    A((void**)&enc_[l], px * ch_[l] * 2);
    A((void**)&genc_[l], px * ch_[l] * 2);
    A((void**)&gin_[l], px * 64 * 2);
    A((void**)&gpre_[l], px * 64 * 2);
    if (l > 0) {
      A((void**)&pin_[l], px * ch_[l - 1] * 2);
      A((void**)&dout_[l], px * ch_[l - 1] * 2);
      A((void**)&gtmp_[l], px * ch_[l - 1] * 2);
    }
    if (l < kLevels - 1) A((void**)&din_[l], px * ch_[l] * 2);
  }
  A((void**)&yb_, size_t(nx) * ny * 4);
  A((void**)&qu_, size_t(nx) * ny * 4);
  A((void**)&qt_, size_t(nx) * ny * 4);
  A((void**)&gy_, size_t(nx) * ny * 4);
  A((void**)&part_, 65536 * 8);
  HG_CHECK(hipHostMalloc((void**)&loss_h_, 65536 * 8));
}

UNet::~UNet() {
  std::vector<void*> ps = {w_, g_, m_, v_, wf_, wb_, wd0_, gtie_, yb_, qu_, qt_, gy_, part_};
  for (int l = 0; l < kLevels; ++l)
    for (void* p : {(void*)enc_[l], (void*)genc_[l], (void*)gin_[l], (void*)gpre_[l],
                    (void*)pin_[l], (void*)dout_[l], (void*)gtmp_[l], (void*)din_[l]})
      ps.push_back(p);
  for (void* p : ps)
    if (p) (void)hipFree(p);
  if (loss_h_) (void)hipHostFree(loss_h_);
}

void UNet::ConvertWeights(hipStream_t st) {
  for (const Layer& L : layer_) {
    const int n = 9 * L.cin * L.cout;
    KConvertW<<<Blk(n), 256, 0, st>>>(L.cin, L.cout, w_ + L.off, wf_ + L.off, wb_ + L.off);
  }
  KFlipE0<<<1, 256, 0, st>>>(w_ + layer_[0].off, wd0_);
}

void UNet::SetWeights(const float* host, hipStream_t st) {
  HG_CHECK(hipMemcpyAsync(w_, host, size_t(nw_) * 4, hipMemcpyHostToDevice, st));
  HG_CHECK(hipMemsetAsync(m_, 0, size_t(nw_) * 4, st));
  HG_CHECK(hipMemsetAsync(v_, 0, size_t(nw_) * 4, st));
  adam_t_ = 0;
  ConvertWeights(st);
  HG_CHECK(hipStreamSynchronize(st));
}

void UNet::GetWeights(float* host, hipStream_t st) {
  HG_CHECK(hipMemcpyAsync(host, w_, size_t(nw_) * 4, hipMemcpyDeviceToHost, st));
  HG_CHECK(hipStreamSynchronize(st));
}

void UNet::InitWeights(uint64_t seed, hipStream_t st) {
  std::mt19937_64 rng(seed);
  std::vector<float> w(nw_, 0.f);
  for (int k = 0; k < 2 * kLevels; ++k) {
    const Layer& L = layer_[k];
    const bool last = (k == 2 * kLevels - 1);
    std::normal_distribution<float> nd(0.f, std::sqrt(2.f / (9.f * L.cin)));
    for (int q = 0; q < 9 * L.cin * L.cout; ++q) w[L.off + q] = last ? 0.f : nd(rng);
  }
  SetWeights(w.data(), st);
}

void UNet::Forward(const float* x, float* y, const float* base, hipStream_t st, const int* done) {
  // encoder
  const bool relu = !linear_;
  KConvE0<<<Blk(long(lw_[0]) * lh_[0]), 256, 0, st>>>(lw_[0], lh_[0], x, w_ + layer_[0].off,
                                                      enc_[0], relu, done);
  for (int l = 1; l < kLevels; ++l) {
    const int W = lw_[l], H = lh_[l];
    KPool<<<Blk(long(W) * H * ch_[l - 1]), 256, 0, st>>>(lw_[l - 1], lh_[l - 1], ch_[l - 1],
                                                        enc_[l - 1], pin_[l], done);
    const Layer& L = layer_[l];
    ConvFwd(L.cin, L.cout, relu, W, H, pin_[l], wf_ + L.off, enc_[l], st, done);
  }
  // decoder: D3 on enc3, then up + skip + conv down to level 1; tied
  // decoder layers are the transposes of the encoder layers (the encoder's
  // dgrad-form weights)
  for (int l = kLevels - 1; l >= 1; --l) {
    const Layer& L = layer_[kLevels + (kLevels - 1 - l)];
    const int W = lw_[l], H = lh_[l];
    const bf* src = (l == kLevels - 1) ? enc_[l] : din_[l];
    if (l < kLevels - 1)
      KUpAdd<<<Blk(long(W) * H * ch_[l]), 256, 0, st>>>(W, H, ch_[l], dout_[l + 1], enc_[l],
                                                        din_[l], done);
    const bf* wt = tied_ ? wb_ + layer_[l].off : wf_ + L.off;
    ConvFwd(L.cin, L.cout, relu, W, H, src, wt, dout_[l], st, done);
  }
  KUpAdd<<<Blk(long(lw_[0]) * lh_[0] * ch_[0]), 256, 0, st>>>(lw_[0], lh_[0], ch_[0], dout_[1],
                                                              enc_[0], din_[0], done);
  KConvD0<<<Blk(long(lw_[0]) * lh_[0]), 256, 0, st>>>(lw_[0], lh_[0], din_[0], D0Weights(), base,
                                                      y, done);
}

const float* UNet::D0Weights() const {
  return tied_ ? wd0_ : w_ + layer_[2 * kLevels - 1].off;
}

void UNet::ZeroGrad(hipStream_t st) { HG_CHECK(hipMemsetAsync(g_, 0, size_t(nw_) * 4, st)); }

void UNet::GetGrad(float* host, hipStream_t st) {
  HG_CHECK(hipMemcpyAsync(host, g_, size_t(nw_) * 4, hipMemcpyDeviceToHost, st));
  HG_CHECK(hipStreamSynchronize(st));
}

void UNet::Backward(const float* x, const float* gy, hipStream_t st) {
  const int W0 = lw_[0], H0 = lh_[0];
  const long n0 = long(W0) * H0;
  // D0: dgrad into gin0 (grad of din0), wgrad
  const Layer& D0 = layer_[2 * kLevels - 1];
  KConvD0Dgrad<<<Blk(n0), 256, 0, st>>>(W0, H0, gy, D0Weights(), gin_[0]);
  if (tied_) {
    HG_CHECK(hipMemsetAsync(gtie_, 0, 9 * 16 * 4, st));
    KEndWgrad<false><<<256, 256, 0, st>>>(W0, H0, nullptr, nullptr, din_[0], gy, gtie_);
    KFoldTied<<<1, 256, 0, st>>>(1, ch_[0], gtie_, g_ + layer_[0].off);
  } else {
    KEndWgrad<false><<<256, 256, 0, st>>>(W0, H0, nullptr, nullptr, din_[0], gy, g_ + D0.off);
  }
  // din0 = up(dout1) + enc0: grad to enc0 and (summed over children) to dout1
  HG_CHECK(hipMemcpyAsync(genc_[0], gin_[0], size_t(n0) * ch_[0] * 2, hipMemcpyDeviceToDevice, st));
  KUpBwd<<<Blk(long(lw_[1]) * lh_[1] * ch_[0]), 256, 0, st>>>(lw_[1], lh_[1], ch_[0], gin_[0],
                                                              gtmp_[1]);
  // decoder D1 .. D_{L-1}
  for (int l = 1; l < kLevels; ++l) {
    const Layer& L = layer_[kLevels + (kLevels - 1 - l)];
    const int W = lw_[l], H = lh_[l];
    const long np = long(W) * H;
    const bf* gp = gtmp_[l];
    if (!linear_) {
      KMask<<<Blk(np * L.cout), 256, 0, st>>>(np * L.cout, gtmp_[l], dout_[l], gpre_[l]);
      gp = gpre_[l];
    }
    const bf* src = (l == kLevels - 1) ? enc_[l] : din_[l];
    if (tied_) {
      const Layer& E = layer_[l];
      HG_CHECK(hipMemsetAsync(gtie_, 0, size_t(9) * E.cin * E.cout * 4, st));
      ConvWgrad(L.cin, L.cout, W, H, src, gp, gtie_, st);
      KFoldTied<<<Blk(9L * E.cin * E.cout), 256, 0, st>>>(E.cin, E.cout, gtie_, g_ + E.off);
      ConvFwd(L.cout, L.cin, false, W, H, gp, wf_ + E.off, gin_[l], st, nullptr);
    } else {
      ConvWgrad(L.cin, L.cout, W, H, src, gp, g_ + L.off, st);
      ConvFwd(L.cout, L.cin, false, W, H, gp, wb_ + L.off, gin_[l], st, nullptr);
    }
    HG_CHECK(hipMemcpyAsync(genc_[l], gin_[l], size_t(np) * ch_[l] * 2, hipMemcpyDeviceToDevice,
                            st));
    if (l < kLevels - 1)
      KUpBwd<<<Blk(long(lw_[l + 1]) * lh_[l + 1] * ch_[l]), 256, 0, st>>>(
          lw_[l + 1], lh_[l + 1], ch_[l], gin_[l], gtmp_[l + 1]);
  }
  // encoder from the top
  for (int l = kLevels - 1; l >= 1; --l) {
    const Layer& L = layer_[l];
    const int W = lw_[l], H = lh_[l];
    const long np = long(W) * H;
    const bf* gp = genc_[l];
    if (!linear_) {
      KMask<<<Blk(np * L.cout), 256, 0, st>>>(np * L.cout, genc_[l], enc_[l], gpre_[l]);
      gp = gpre_[l];
    }
    ConvWgrad(L.cin, L.cout, W, H, pin_[l], gp, g_ + L.off, st);
    ConvFwd(L.cout, L.cin, false, W, H, gp, wb_ + L.off, gin_[l], st, nullptr);
    KPoolBwdAdd<<<Blk(long(lw_[l - 1]) * lh_[l - 1] * ch_[l - 1]), 256, 0, st>>>(
        lw_[l - 1], lh_[l - 1], ch_[l - 1], gin_[l], genc_[l - 1]);
  }
  const bf* gp0 = genc_[0];
  if (!linear_) {
    KMask<<<Blk(n0 * ch_[0]), 256, 0, st>>>(n0 * ch_[0], genc_[0], enc_[0], gpre_[0]);
    gp0 = gpre_[0];
  }
  KEndWgrad<true><<<256, 256, 0, st>>>(W0, H0, x, gp0, nullptr, nullptr, g_ + layer_[0].off);
}

void UNet::AdamStep(float lr, hipStream_t st) {
  ++adam_t_;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, float(adam_t_)), c2 = 1.f - std::pow(b2, float(adam_t_));
  KAdam<<<Blk(nw_), 256, 0, st>>>(nw_, w_, g_, m_, v_, lr, b1, b2, c1, c2);
  ConvertWeights(st);
}

void UNet::TrainStep(const float* x, const float* target, float lr, hipStream_t st,
                     double* loss) {
  Forward(x, yb_, nullptr, st, nullptr);
  const long n = long(nx_) * ny_;
  const int nb = 1024;
  if (mkx_ && scale_free_) {
    KCosDots<<<nb, 256, 0, st>>>(nx_, ny_, mkx_, mky_, yb_, oscale_, target, qu_, qt_, part_);
    KCosGrad<<<nb, 256, 0, st>>>(n, nb, part_, oscale_, qu_, qt_, gy_, part_ + 3 * nb);
    if (loss) HG_CHECK(hipMemcpyAsync(loss_h_, part_ + 3 * nb, 16, hipMemcpyDeviceToHost, st));
  } else {
    if (mkx_) {
      KDiff<<<nb, 256, 0, st>>>(n, yb_, target, oscale_);
      KEnergyLoss<<<nb, 256, 0, st>>>(nx_, ny_, mkx_, mky_, yb_, oscale_, gy_, part_);
    } else {
      KLoss<<<nb, 256, 0, st>>>(n, yb_, target, gy_, part_);
    }
    if (loss) HG_CHECK(hipMemcpyAsync(loss_h_, part_, nb * 8, hipMemcpyDeviceToHost, st));
  }
  ZeroGrad(st);
  Backward(x, gy_, st);
  AdamStep(lr, st);
  if (loss) {
    HG_CHECK(hipStreamSynchronize(st));
    if (mkx_ && scale_free_) {
      *loss = loss_h_[0];
      last_alpha_ = loss_h_[1];
    } else {
      double s = 0.;
      for (int b = 0; b < nb; ++b) s += loss_h_[b];
      *loss = s / double(n);
    }
  }
}

void UNet::ScaleOutput(float a, hipStream_t st) {
  if (!(a > 0.f)) throw std::runtime_error("UNet::ScaleOutput: factor must be positive");
  // the tied network is quadratic in the first layer, the untied one linear
  // in the last
  const Layer& L = tied_ ? layer_[0] : layer_[2 * kLevels - 1];
  const int n = 9 * L.cin * L.cout;
  KScaleW<<<Blk(n), 256, 0, st>>>(n, w_ + L.off, tied_ ? std::sqrt(a) : a);
  ConvertWeights(st);
}

double UNet::ForwardFlops() const {
  double f = 0.;
  for (const Layer& L : layer_)
    f += 18. * L.cin * L.cout * double(lw_[L.level]) * lh_[L.level];
  return f;
}

double UNet::MfmaForwardFlops() const {
  double f = 0.;
  for (const Layer& L : layer_)
    if (L.cin % 16 == 0 && L.cout % 16 == 0)
      f += 18. * L.cin * L.cout * double(lw_[L.level]) * lh_[L.level];
  return f;
}

// forward + weight gradients of every layer + input gradients of every
// layer but the first
double UNet::TrainFlops() const {
  double f = 2. * ForwardFlops();
  const Layer& E0 = layer_[0];
  f += ForwardFlops() - 18. * E0.cin * E0.cout * double(lw_[0]) * lh_[0];
  return f;
}

}  // namespace hg
