// unet.hip -- kernels of the learned warm start: a 4-level U-Net (3x3
// convolutions, average pooling, nearest-neighbour upsampling, skip
// concatenation) with BF16 weights and FP32 accumulation on gfx950 MFMA.
//
// Tensors are NHWC FP32 in HBM; weights are [ky][kx][ci][co] (FP32 master and
// a BF16 copy the kernels read).  Every 3x3 convolution and both of its
// gradients are one implicit GEMM on v_mfma_f32_16x16x32_bf16:
//   forward       C[p][co] = sum_k A[p][k] B[k][co],  A[p][(t,ci)] = in(p + t - 1), B = W
//   input grad    C[p][ci] = sum_k A[p][k] B[k][ci],  A[p][(t,co)] = dout(p - t + 1), B[(t,co)][ci] = W[t][ci][co]
//   weight grad   C[(t,ci)][co] = sum_p A[(t,ci)][p] dout[p][co]   (pixel range split over workgroups)
// A and B tiles are gathered (zero padding, two concatenated inputs),
// rounded to BF16 into LDS and multiplied on the matrix cores.
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <algorithm>
#include <cstdint>

#include "../../include/mfista_learn.h"

namespace unet {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int BM = 64, BN = 64, BK = 32, NT = 256;

__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even, as the MFMA inputs
__device__ inline uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

enum Mode { kFwd = 0, kBwdData = 1, kBwdWeight = 2 };

struct ConvArgs {
  const float* in1;  // forward / weight grad: the layer input (first C1 channels)
  const float* in2;  // its second part (C2 channels, may be null)
  int C1, C2;
  const float* dout;  // input / weight grad: gradient of the layer output (Cout channels)
  const uint16_t* w;  // BF16 weights [9][Cin][Cout]
  const float* bias;
  float* out1;  // forward: output; input grad: gradient of in1
  float* out2;  // input grad: gradient of in2
  float* dw;    // weight grad: accumulated (FP32 atomics)
  int B, H, W, Cout, relu;
  int pix_per_split;  // weight grad: pixels per workgroup along grid.z
};

// element (r, c) of the GEMM operand A (rows x K) and B (K x cols)
template <int MODE>
__device__ inline float ga(const ConvArgs& a, long r, long c) {
  const int Cin = a.C1 + a.C2;
  if (MODE == kFwd || MODE == kBwdWeight) {
    // forward: r = pixel, c = (tap, ci); weight grad: r = (tap, ci), c = pixel
    const long p = (MODE == kFwd) ? r : c;
    const long k = (MODE == kFwd) ? c : r;
    const int t = (int)(k / Cin), ci = (int)(k - (long)t * Cin);
    if (t >= 9) return 0.f;
    const int x = (int)(p % a.W);
    const long q = p / a.W;
    const int y = (int)(q % a.H);
    const int b = (int)(q / a.H);
    if (b >= a.B) return 0.f;
    const int yy = y + t / 3 - 1, xx = x + t % 3 - 1;
    if (yy < 0 || yy >= a.H || xx < 0 || xx >= a.W) return 0.f;
    const long pp = ((long)b * a.H + yy) * a.W + xx;
    return (ci < a.C1) ? a.in1[pp * a.C1 + ci] : a.in2[pp * a.C2 + (ci - a.C1)];
  } else {
    // input grad: r = pixel, c = (tap, co): dout at p - tap + 1
    const int t = (int)(c / a.Cout), co = (int)(c - (long)t * a.Cout);
    if (t >= 9) return 0.f;
    const int x = (int)(r % a.W);
    const long q = r / a.W;
    const int y = (int)(q % a.H);
    const int b = (int)(q / a.H);
    if (b >= a.B) return 0.f;
    const int yy = y - (t / 3 - 1), xx = x - (t % 3 - 1);
    if (yy < 0 || yy >= a.H || xx < 0 || xx >= a.W) return 0.f;
    return a.dout[(((long)b * a.H + yy) * a.W + xx) * a.Cout + co];
  }
}

template <int MODE>
__device__ inline float gb(const ConvArgs& a, long k, int n, long K, int N) {
  if (k >= K || n >= N) return 0.f;
  const int Cin = a.C1 + a.C2;
  if (MODE == kFwd) return bf2f(a.w[k * a.Cout + n]);  // B[(t,ci)][co]
  if (MODE == kBwdData) {                               // B[(t,co)][ci] = W[t][ci][co]
    const int t = (int)(k / a.Cout), co = (int)(k - (long)t * a.Cout);
    return bf2f(a.w[((long)t * Cin + n) * a.Cout + co]);
  }
  return a.dout[k * a.Cout + n];  // weight grad: B[p][co]
}

template <int MODE>
__global__ void __launch_bounds__(NT) k_conv_gemm(ConvArgs a, long M, int N, long K) {
  __shared__ uint16_t As[BM][BK + 8];  // [m][k], padded rows
  __shared__ uint16_t Bs[BN][BK + 8];  // [n][k] (transposed)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;  // 32 x 32 sub-tile of the wave
  const long m0 = (long)blockIdx.x * BM;
  const int n0 = blockIdx.y * BN;
  long k_begin = 0, k_end = K;
  if (MODE == kBwdWeight) {  // split the pixel reduction
    k_begin = (long)blockIdx.z * a.pix_per_split;
    k_end = k_begin + a.pix_per_split;
    if (k_end > K) k_end = K;
  }
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (long k0 = k_begin; k0 < k_end; k0 += BK) {
    // stage A (BM x BK) and B (BK x BN) as BF16; 8 elements of each per thread
#pragma unroll
    for (int e = 0; e < (BM * BK) / NT; e++) {
      const int idx = tid + e * NT;
      const int mm = idx / BK, kk = idx % BK;
      const long gm = m0 + mm, gk = k0 + kk;
      float v = 0.f;
      if (gm < M && gk < k_end) v = ga<MODE>(a, gm, gk);
      As[mm][kk] = f2bf(v);
    }
#pragma unroll
    for (int e = 0; e < (BN * BK) / NT; e++) {
      const int idx = tid + e * NT;
      const int nn = idx / BK, kk = idx % BK;
      const long gk = k0 + kk;
      float v = 0.f;
      if (gk < k_end) v = gb<MODE>(a, gk, n0 + nn, K, N);
      Bs[nn][kk] = f2bf(v);
    }
    __syncthreads();
    // fragments: lane l holds A[row l % 16][k (l / 16) * 8 .. +7] and B[k ...][col l % 16]
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const uint16_t* pa = &As[wm * 32 + i * 16 + (lane & 15)][(lane >> 4) * 8];
      const uint16_t* pb = &Bs[wn * 32 + i * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        af[i][q] = __builtin_bit_cast(__bf16, pa[q]);
        bfr[i][q] = __builtin_bit_cast(__bf16, pb[q]);
      }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // result: lane l holds C[row (l / 16) * 4 + q][col l % 16] of each 16 x 16 block
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const long gm = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + q;
        const int gn = n0 + wn * 32 + j * 16 + (lane & 15);
        if (gm >= M || gn >= N) continue;
        float v = acc[i][j][q];
        if (MODE == kFwd) {
          if (a.bias) v += a.bias[gn];
          if (a.relu && v < 0.f) v = 0.f;
          a.out1[gm * a.Cout + gn] = v;
        } else if (MODE == kBwdData) {
          if (gn < a.C1) {
            if (a.out1) a.out1[gm * a.C1 + gn] = v;
          } else if (a.out2) {
            a.out2[gm * a.C2 + (gn - a.C1)] = v;
          }
        } else {
          unsafeAtomicAdd(&a.dw[gm * a.Cout + gn], v);
        }
      }
}

// Fast path of k_conv_gemm when every channel count involved is a multiple
// of 32 (all layers but the first): a 32-wide K slice then lies in one tap
// and one input tensor, so each thread stages 8 consecutive channels of one
// pixel with two float4 loads and one address computation.
struct PixelIdx {
  int b, y, x;
};

__device__ inline PixelIdx pix(long p, int H, int W) {
  PixelIdx r;
  r.x = (int)(p % W);
  const long q = p / W;
  r.y = (int)(q % H);
  r.b = (int)(q / H);
  return r;
}

// 8 consecutive channels c0 .. c0+7 of `src` (C channels) at pixel (b, yy, xx), zero outside
__device__ inline void load8(const float* src, int C, int c0, int b, int yy, int xx, int H, int W, int B,
                             float (&v)[8]) {
  if (b >= B || yy < 0 || yy >= H || xx < 0 || xx >= W) {
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = 0.f;
    return;
  }
  const float4* q = reinterpret_cast<const float4*>(src + (((long)b * H + yy) * W + xx) * C + c0);
  const float4 u0 = q[0], u1 = q[1];
  v[0] = u0.x, v[1] = u0.y, v[2] = u0.z, v[3] = u0.w, v[4] = u1.x, v[5] = u1.y, v[6] = u1.z, v[7] = u1.w;
}

template <int MODE>
__global__ void __launch_bounds__(NT) k_conv_gemm_vec(ConvArgs a, long M, int N, long K) {
  __shared__ uint16_t As[BM][BK + 8];
  __shared__ uint16_t Bs[BN][BK + 8];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const long m0 = (long)blockIdx.x * BM;
  const int n0 = blockIdx.y * BN;
  const int Cin = a.C1 + a.C2;
  long k_begin = 0, k_end = K;
  if (MODE == kBwdWeight) {
    k_begin = (long)blockIdx.z * a.pix_per_split;
    k_end = k_begin + a.pix_per_split;
    if (k_end > K) k_end = K;
  }
  floatx4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // per-thread staging coordinates that do not change along K
  PixelIdx pa{0, 0, 0};
  long gma = 0;
  if (MODE != kBwdWeight) {  // A rows are pixels: thread -> (row tid / 4, channels (tid % 4) * 8)
    gma = m0 + tid / 4;
    pa = pix(gma, a.H, a.W);
  }
  for (long k0 = k_begin; k0 < k_end; k0 += BK) {
    float v[8];
    if (MODE == kFwd) {
      const int t = (int)(k0 / Cin), c0 = (int)(k0 - (long)t * Cin) + (tid % 4) * 8;
      const int yy = pa.y + t / 3 - 1, xx = pa.x + t % 3 - 1;
      const int b = (gma < M) ? pa.b : a.B;
      if (c0 < a.C1)
        load8(a.in1, a.C1, c0, b, yy, xx, a.H, a.W, a.B, v);
      else
        load8(a.in2, a.C2, c0 - a.C1, b, yy, xx, a.H, a.W, a.B, v);
      uint16_t* dst = &As[tid / 4][(tid % 4) * 8];
#pragma unroll
      for (int i = 0; i < 8; i++) dst[i] = f2bf(v[i]);
      // B[k][co] = W[k][co]: row kk = tid / 8, columns (tid % 8) * 8 .. +7
      const int kk = tid / 8, nn = (tid % 8) * 8;
      const uint16_t* w = a.w + (k0 + kk) * a.Cout + n0 + nn;
      if (n0 + nn + 8 <= N) {
        const uint4 u = *reinterpret_cast<const uint4*>(w);
        const uint16_t* h = reinterpret_cast<const uint16_t*>(&u);
#pragma unroll
        for (int i = 0; i < 8; i++) Bs[nn + i][kk] = h[i];
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) Bs[nn + i][kk] = (n0 + nn + i < N) ? w[i] : (uint16_t)0;
      }
    } else if (MODE == kBwdData) {
      const int t = (int)(k0 / a.Cout), c0 = (int)(k0 - (long)t * a.Cout) + (tid % 4) * 8;
      const int yy = pa.y - (t / 3 - 1), xx = pa.x - (t % 3 - 1);
      const int b = (gma < M) ? pa.b : a.B;
      load8(a.dout, a.Cout, c0, b, yy, xx, a.H, a.W, a.B, v);
      uint16_t* dst = &As[tid / 4][(tid % 4) * 8];
#pragma unroll
      for (int i = 0; i < 8; i++) dst[i] = f2bf(v[i]);
      // B[(t,co)][ci] = W[t][ci][co]: row n = tid / 4 (ci), co (tid % 4) * 8 .. +7 contiguous
      const int nn = tid / 4, cc = (int)(k0 - (long)t * a.Cout) + (tid % 4) * 8;
      uint16_t* bd = &Bs[nn][(tid % 4) * 8];
      if (n0 + nn < N) {
        const uint4 u = *reinterpret_cast<const uint4*>(a.w + ((long)t * Cin + n0 + nn) * a.Cout + cc);
        const uint16_t* h = reinterpret_cast<const uint16_t*>(&u);
#pragma unroll
        for (int i = 0; i < 8; i++) bd[i] = h[i];
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++) bd[i] = 0;
      }
    } else {
      // A[(t,ci)][p]: pixel kk = tid % 32, rows (tid / 32) * 8 .. +7 (one tap, 8 channels)
      const int kk = tid % 32, mm = (tid / 32) * 8;
      const long gp = k0 + kk;
      const long gm = m0 + mm;
      const int t = (int)(gm / Cin), c0 = (int)(gm - (long)t * Cin);
      PixelIdx q = pix(gp, a.H, a.W);
      const int b = (gp < k_end && t < 9) ? q.b : a.B;
      const int yy = q.y + t / 3 - 1, xx = q.x + t % 3 - 1;
      if (c0 < a.C1)
        load8(a.in1, a.C1, c0, b, yy, xx, a.H, a.W, a.B, v);
      else
        load8(a.in2, a.C2, c0 - a.C1, b, yy, xx, a.H, a.W, a.B, v);
#pragma unroll
      for (int i = 0; i < 8; i++) As[mm + i][kk] = f2bf(v[i]);
      // B[p][co] = dout[p][co]: 8 consecutive co of pixel kk
      const int nn = (tid / 32) * 8;
      if (gp < k_end && n0 + nn + 8 <= N) {
        const float4* d = reinterpret_cast<const float4*>(a.dout + gp * a.Cout + n0 + nn);
        const float4 u0 = d[0], u1 = d[1];
        const float w8[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
        for (int i = 0; i < 8; i++) Bs[nn + i][kk] = f2bf(w8[i]);
      } else {
#pragma unroll
        for (int i = 0; i < 8; i++)
          Bs[nn + i][kk] = (gp < k_end && n0 + nn + i < N) ? f2bf(a.dout[gp * a.Cout + n0 + nn + i]) : (uint16_t)0;
      }
    }
    __syncthreads();
    bf16x8 af[2], bfr[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
      const uint16_t* qa = &As[wm * 32 + i * 16 + (lane & 15)][(lane >> 4) * 8];
      const uint16_t* qb = &Bs[wn * 32 + i * 16 + (lane & 15)][(lane >> 4) * 8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        af[i][q] = __builtin_bit_cast(__bf16, qa[q]);
        bfr[i][q] = __builtin_bit_cast(__bf16, qb[q]);
      }
    }
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const long gm = m0 + wm * 32 + i * 16 + (lane >> 4) * 4 + q;
        const int gn = n0 + wn * 32 + j * 16 + (lane & 15);
        if (gm >= M || gn >= N) continue;
        float v = acc[i][j][q];
        if (MODE == kFwd) {
          if (a.bias) v += a.bias[gn];
          if (a.relu && v < 0.f) v = 0.f;
          a.out1[gm * a.Cout + gn] = v;
        } else if (MODE == kBwdData) {
          if (gn < a.C1) {
            if (a.out1) a.out1[gm * a.C1 + gn] = v;
          } else if (a.out2) {
            a.out2[gm * a.C2 + (gn - a.C1)] = v;
          }
        } else {
          unsafeAtomicAdd(&a.dw[gm * a.Cout + gn], v);
        }
      }
}

// bias gradient: db[co] += sum over pixels of dout[p][co]
__global__ void k_bias_grad(const float* __restrict__ dout, float* __restrict__ db, long P, int C) {
  const int c = blockIdx.y * 64 + (threadIdx.x & 63);
  const int sub = threadIdx.x >> 6;
  if (c >= C) return;
  float s = 0.f;
  for (long p = (long)blockIdx.x * 4 + sub; p < P; p += (long)gridDim.x * 4) s += dout[p * C + c];
  unsafeAtomicAdd(&db[c], s);
}

__global__ void k_relu_bwd(const float* __restrict__ out, float* __restrict__ g, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    if (out[i] <= 0.f) g[i] = 0.f;
}

// 2x2 average pooling, NHWC; H, W of the input
__global__ void k_pool_fwd(const float* __restrict__ in, float* __restrict__ out, int B, int H, int W, int C) {
  const int Ho = H / 2, Wo = W / 2;
  const long n = (long)B * Ho * Wo * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long q = i / C;
    const int x = (int)(q % Wo);
    q /= Wo;
    const int y = (int)(q % Ho);
    const int b = (int)(q / Ho);
    const long base = (((long)b * H + 2 * y) * W + 2 * x) * C + c;
    out[i] = 0.25f * (in[base] + in[base + C] + in[base + (long)W * C] + in[base + (long)W * C + C]);
  }
}

__global__ void k_pool_bwd(const float* __restrict__ dout, float* __restrict__ din, int B, int H, int W, int C,
                           int acc) {
  const long n = (long)B * H * W * C;
  const int Wo = W / 2, Ho = H / 2;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long q = i / C;
    const int x = (int)(q % W);
    q /= W;
    const int y = (int)(q % H);
    const int b = (int)(q / H);
    const float v = 0.25f * dout[(((long)b * Ho + y / 2) * Wo + x / 2) * C + c];
    din[i] = acc ? din[i] + v : v;
  }
}

// nearest-neighbour 2x upsampling; H, W of the input
__global__ void k_up_fwd(const float* __restrict__ in, float* __restrict__ out, int B, int H, int W, int C) {
  const int Ho = 2 * H, Wo = 2 * W;
  const long n = (long)B * Ho * Wo * C;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long q = i / C;
    const int x = (int)(q % Wo);
    q /= Wo;
    const int y = (int)(q % Ho);
    const int b = (int)(q / Ho);
    out[i] = in[(((long)b * H + y / 2) * W + x / 2) * C + c];
  }
}

__global__ void k_up_bwd(const float* __restrict__ dout, float* __restrict__ din, int B, int H, int W, int C) {
  const long n = (long)B * H * W * C;
  const int Wo = 2 * W;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const int c = (int)(i % C);
    long q = i / C;
    const int x = (int)(q % W);
    q /= W;
    const int y = (int)(q % H);
    const int b = (int)(q / H);
    const long o = (((long)b * 2 * H + 2 * y) * Wo + 2 * x) * C + c;
    din[i] = dout[o] + dout[o + C] + dout[o + (long)Wo * C] + dout[o + (long)Wo * C + C];
  }
}

// 1x1 convolution to one channel (the head), BF16 weights
__global__ void k_head_fwd(const float* __restrict__ in, const uint16_t* __restrict__ w, const float* __restrict__ bias,
                           float* __restrict__ out, long P, int C, int relu) {
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < P; p += (long)gridDim.x * blockDim.x) {
    float s = bias[0];
    for (int c = 0; c < C; c++) s += in[p * C + c] * bf2f(w[c]);
    out[p] = (relu && s < 0.f) ? 0.f : s;
  }
}

// head backward: din[p][c] = dout[p] w[c]; dw[c] += sum_p dout[p] in[p][c]; db += sum_p dout[p]
__global__ void k_head_bwd(const float* __restrict__ in, const uint16_t* __restrict__ w, const float* __restrict__ dout,
                           float* __restrict__ din, float* __restrict__ dw, float* __restrict__ db, long P, int C) {
  __shared__ float red_w[8][32];
  __shared__ float red_b[8];
  const int c = threadIdx.x & 31;  // C <= 32
  const int sub = threadIdx.x >> 5;
  float sw = 0.f, sb = 0.f;
  for (long p = (long)blockIdx.x * 8 + sub; p < P; p += (long)gridDim.x * 8) {
    const float g = dout[p];
    if (c < C) {
      if (din) din[p * C + c] = g * bf2f(w[c]);
      sw += g * in[p * C + c];
    }
    sb += g;
  }
  red_w[sub][c] = sw;
  if (c == 0) red_b[sub] = sb;
  __syncthreads();
  if (threadIdx.x < 32 && c < C) {
    float s = 0.f;
    for (int q = 0; q < 8; q++) s += red_w[q][c];
    unsafeAtomicAdd(&dw[c], s);
  }
  if (threadIdx.x == 0) {
    float s = 0.f;
    for (int q = 0; q < 8; q++) s += red_b[q];
    unsafeAtomicAdd(db, s);
  }
}

// mean squared error over n elements: loss (FP64 partials) and dy = 2 (y - t) / n
__global__ void k_mse(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dy,
                      double* __restrict__ loss, long n) {
  __shared__ double red[NT];
  double s = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float d = y[i] - t[i];
    s += (double)d * d;
    if (dy) dy[i] = 2.f * d / (float)n;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] / (double)n);
}

// Per-sample weighted squared error: loss = sum_b wt[b] |y_b - t_b|^2 / n and
// its gradient; sample b is elements [b per, (b + 1) per)
__global__ void k_mse_w(const float* __restrict__ y, const float* __restrict__ t, const float* __restrict__ wt,
                        long per, float* __restrict__ dy, double* __restrict__ loss, long n) {
  __shared__ double red[NT];
  double s = 0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float w = wt[i / per];
    const float d = y[i] - t[i];
    s += (double)w * d * d;
    if (dy) dy[i] = 2.f * w * d / (float)n;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = NT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] / (double)n);
}

// Adam on the FP32 master weights; refreshes the BF16 copy
__global__ void k_adam(float* __restrict__ w, uint16_t* __restrict__ w16, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, long n, float lr, float b1, float b2,
                       float eps, float c1, float c2) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    w16[i] = f2bf(wi);
  }
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ w16, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    w16[i] = f2bf(w[i]);
}

static inline int grid1(long n) {
  long b = (n + NT - 1) / NT;
  if (b > 65535) b = 65535;
  return (int)(b < 1 ? 1 : b);
}

}  // namespace unet

using namespace unet;

#define UN_CHECK(x)                                  \
  do {                                               \
    if ((x) != hipSuccess) return UNET_E_HIP;        \
  } while (0)

static inline bool al16(const void* p) { return ((uintptr_t)p & 15) == 0; }

extern "C" {

int unet_conv3x3_fwd(const float* in1, int C1, const float* in2, int C2, const uint16_t* w, const float* bias,
                     float* out, int B, int H, int W, int Cout, int relu, void* stream) {
  if (!in1 || !w || !out || C1 <= 0 || C2 < 0 || (C2 > 0 && !in2) || B <= 0 || H <= 0 || W <= 0 || Cout <= 0)
    return UNET_E_ARG;
  ConvArgs a{};
  a.in1 = in1;
  a.in2 = in2;
  a.C1 = C1;
  a.C2 = C2;
  a.w = w;
  a.bias = bias;
  a.out1 = out;
  a.B = B;
  a.H = H;
  a.W = W;
  a.Cout = Cout;
  a.relu = relu;
  const long M = (long)B * H * W, K = 9L * (C1 + C2);
  dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((Cout + BN - 1) / BN));
  if (C1 % 32 == 0 && C2 % 32 == 0 && Cout % 8 == 0 && al16(in1) && al16(in2) && al16(w))
    hipLaunchKernelGGL(k_conv_gemm_vec<kFwd>, grid, dim3(NT), 0, (hipStream_t)stream, a, M, Cout, K);
  else
    hipLaunchKernelGGL(k_conv_gemm<kFwd>, grid, dim3(NT), 0, (hipStream_t)stream, a, M, Cout, K);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_conv3x3_bwd_data(const float* dout, int Cout, const uint16_t* w, float* din1, int C1, float* din2, int C2,
                          int B, int H, int W, void* stream) {
  if (!dout || !w || C1 <= 0 || C2 < 0 || B <= 0 || H <= 0 || W <= 0 || Cout <= 0) return UNET_E_ARG;
  ConvArgs a{};
  a.dout = dout;
  a.w = w;
  a.C1 = C1;
  a.C2 = C2;
  a.out1 = din1;
  a.out2 = din2;
  a.B = B;
  a.H = H;
  a.W = W;
  a.Cout = Cout;
  const long M = (long)B * H * W, K = 9L * Cout;
  const int N = C1 + C2;
  dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((N + BN - 1) / BN));
  if (Cout % 32 == 0 && al16(dout) && al16(w))
    hipLaunchKernelGGL(k_conv_gemm_vec<kBwdData>, grid, dim3(NT), 0, (hipStream_t)stream, a, M, N, K);
  else
    hipLaunchKernelGGL(k_conv_gemm<kBwdData>, grid, dim3(NT), 0, (hipStream_t)stream, a, M, N, K);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_conv3x3_bwd_weight(const float* in1, int C1, const float* in2, int C2, const float* dout, int Cout,
                            float* dw, float* db, int B, int H, int W, void* stream) {
  if (!in1 || !dout || !dw || C1 <= 0 || C2 < 0 || (C2 > 0 && !in2) || B <= 0 || H <= 0 || W <= 0 || Cout <= 0)
    return UNET_E_ARG;
  ConvArgs a{};
  a.in1 = in1;
  a.in2 = in2;
  a.C1 = C1;
  a.C2 = C2;
  a.dout = dout;
  a.dw = dw;
  a.B = B;
  a.H = H;
  a.W = W;
  a.Cout = Cout;
  const long P = (long)B * H * W;
  const long M = 9L * (C1 + C2);
  // enough workgroups to fill the chip: split the pixel reduction
  const long tiles = ((M + BM - 1) / BM) * ((Cout + BN - 1) / BN);
  long splits = (2048 + tiles - 1) / tiles;
  long per = (P + splits - 1) / splits;
  per = ((per + BK - 1) / BK) * BK;
  splits = (P + per - 1) / per;
  a.pix_per_split = (int)per;
  dim3 grid((unsigned)((M + BM - 1) / BM), (unsigned)((Cout + BN - 1) / BN), (unsigned)splits);
  if (C1 % 32 == 0 && C2 % 32 == 0 && Cout % 8 == 0 && al16(in1) && al16(in2) && al16(dout))
    hipLaunchKernelGGL(k_conv_gemm_vec<kBwdWeight>, grid, dim3(NT), 0, (hipStream_t)stream, a, M, Cout, P);
  else
    hipLaunchKernelGGL(k_conv_gemm<kBwdWeight>, grid, dim3(NT), 0, (hipStream_t)stream, a, M, Cout, P);
  UN_CHECK(hipGetLastError());
  if (db) {
    dim3 g2((unsigned)std::min<long>((P + 3) / 4, 1024), (unsigned)((Cout + 63) / 64));
    hipLaunchKernelGGL(k_bias_grad, g2, dim3(NT), 0, (hipStream_t)stream, dout, db, P, Cout);
    UN_CHECK(hipGetLastError());
  }
  return 0;
}

int unet_relu_bwd(const float* out, float* dout, long n, void* stream) {
  if (!out || !dout || n <= 0) return UNET_E_ARG;
  hipLaunchKernelGGL(k_relu_bwd, dim3(grid1(n)), dim3(NT), 0, (hipStream_t)stream, out, dout, n);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_avgpool2_fwd(const float* in, float* out, int B, int H, int W, int C, void* stream) {
  if (!in || !out || (H & 1) || (W & 1) || B <= 0 || C <= 0) return UNET_E_ARG;
  hipLaunchKernelGGL(k_pool_fwd, dim3(grid1((long)B * H * W * C / 4)), dim3(NT), 0, (hipStream_t)stream, in, out, B,
                     H, W, C);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_avgpool2_bwd(const float* dout, float* din, int B, int H, int W, int C, int accumulate, void* stream) {
  if (!dout || !din || (H & 1) || (W & 1) || B <= 0 || C <= 0) return UNET_E_ARG;
  hipLaunchKernelGGL(k_pool_bwd, dim3(grid1((long)B * H * W * C)), dim3(NT), 0, (hipStream_t)stream, dout, din, B, H,
                     W, C, accumulate);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_upsample2_fwd(const float* in, float* out, int B, int H, int W, int C, void* stream) {
  if (!in || !out || B <= 0 || H <= 0 || W <= 0 || C <= 0) return UNET_E_ARG;
  hipLaunchKernelGGL(k_up_fwd, dim3(grid1((long)B * H * W * C * 4)), dim3(NT), 0, (hipStream_t)stream, in, out, B, H,
                     W, C);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_upsample2_bwd(const float* dout, float* din, int B, int H, int W, int C, void* stream) {
  if (!dout || !din || B <= 0 || H <= 0 || W <= 0 || C <= 0) return UNET_E_ARG;
  hipLaunchKernelGGL(k_up_bwd, dim3(grid1((long)B * H * W * C)), dim3(NT), 0, (hipStream_t)stream, dout, din, B, H,
                     W, C);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_head_fwd(const float* in, int C, const uint16_t* w, const float* bias, float* out, long P, int relu,
                  void* stream) {
  if (!in || !w || !bias || !out || C <= 0 || C > 32 || P <= 0) return UNET_E_ARG;
  hipLaunchKernelGGL(k_head_fwd, dim3(grid1(P)), dim3(NT), 0, (hipStream_t)stream, in, w, bias, out, P, C, relu);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_head_bwd(const float* in, int C, const uint16_t* w, const float* dout, float* din, float* dw, float* db,
                  long P, void* stream) {
  if (!in || !w || !dout || !dw || !db || C <= 0 || C > 32 || P <= 0) return UNET_E_ARG;
  hipLaunchKernelGGL(k_head_bwd, dim3((unsigned)std::min<long>((P + 7) / 8, 2048)), dim3(NT), 0, (hipStream_t)stream,
                     in, w, dout, din, dw, db, P, C);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_mse(const float* y, const float* t, float* dy, double* loss, long n, void* stream) {
  if (!y || !t || !loss || n <= 0) return UNET_E_ARG;
  UN_CHECK(hipMemsetAsync(loss, 0, sizeof(double), (hipStream_t)stream));
  hipLaunchKernelGGL(k_mse, dim3((unsigned)std::min<long>(grid1(n), 1024)), dim3(NT), 0, (hipStream_t)stream, y, t, dy,
                     loss, n);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_mse_weighted(const float* y, const float* t, const float* wt, long per, float* dy, double* loss, long n,
                      void* stream) {
  if (!y || !t || !wt || !loss || n <= 0 || per <= 0 || n % per != 0) return UNET_E_ARG;
  UN_CHECK(hipMemsetAsync(loss, 0, sizeof(double), (hipStream_t)stream));
  hipLaunchKernelGGL(k_mse_w, dim3((unsigned)std::min<long>(grid1(n), 1024)), dim3(NT), 0, (hipStream_t)stream, y, t,
                     wt, per, dy, loss, n);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_adam(float* w, uint16_t* w16, const float* g, float* m, float* v, long n, float lr, float b1, float b2,
              float eps, int step, void* stream) {
  if (!w || !w16 || !g || !m || !v || n <= 0 || step < 1) return UNET_E_ARG;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(NT), 0, (hipStream_t)stream, w, w16, g, m, v, n, lr, b1, b2, eps,
                     c1, c2);
  UN_CHECK(hipGetLastError());
  return 0;
}

int unet_to_bf16(const float* w, uint16_t* w16, long n, void* stream) {
  if (!w || !w16 || n <= 0) return UNET_E_ARG;
  hipLaunchKernelGGL(k_to_bf16, dim3(grid1(n)), dim3(NT), 0, (hipStream_t)stream, w, w16, n);
  UN_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
