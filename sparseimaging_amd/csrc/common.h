// common.h -- shared types and device helpers for the MI355X MFISTA library.
//
// One process, one GPU (gfx950).  Every kernel is written for 64-wide
// wavefronts; reductions are done in fixed order so a solve is bit-for-bit
// reproducible run to run.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define MF_CHECK(expr)                                                                    \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess) {                                                               \
      fprintf(stderr, "[mfista_amd] HIP error %s at %s:%d: %s\n", hipGetErrorName(_e),   \
              __FILE__, __LINE__, #expr);                                                 \
      return -100;                                                                        \
    }                                                                                     \
  } while (0)

#define MF_WAVE 64
#define MF_BLOCK 256

namespace mf {

template <typename R>
struct cplx_t;
template <>
struct cplx_t<float> {
  using type = float2;
};
template <>
struct cplx_t<double> {
  using type = double2;
};
template <typename R>
using cplx = typename cplx_t<R>::type;

template <typename R>
__host__ __device__ inline cplx<R> mk(R a, R b) {
  cplx<R> r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename C>
__host__ __device__ inline C cadd(C a, C b) {
  a.x += b.x;
  a.y += b.y;
  return a;
}
template <typename C>
__host__ __device__ inline C csub(C a, C b) {
  a.x -= b.x;
  a.y -= b.y;
  return a;
}
template <typename C>
__host__ __device__ inline C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
template <typename C>
__host__ __device__ inline C cconj(C a) {
  a.y = -a.y;
  return a;
}
template <typename C, typename R>
__host__ __device__ inline C cscale(C a, R s) {
  a.x *= s;
  a.y *= s;
  return a;
}

__device__ inline float dev_exp(float a) { return __expf(a); }
__device__ inline double dev_exp(double a) { return exp(a); }

// Fixed-order block reduction of NV doubles per thread; thread 0 of the block
// receives the sums.  blockDim.x must be a multiple of 64 and <= 1024.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* smem /* >= 16*NV */) {
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = v[q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    v[q] = s;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if (lane == 0)
    for (int q = 0; q < NV; q++) smem[wid * NV + q] = v[q];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < NV; q++) {
      double s = 0;
      for (int w = 0; w < nw; w++) s += smem[w * NV + q];
      v[q] = s;
    }
  }
  __syncthreads();
}

}  // namespace mf
