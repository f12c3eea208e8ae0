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

// Loads and stores through the global address space.  Pointers that the
// kernels read out of argument structs are generic, so plain dereferences
// compile to FLAT instructions.  A FLAT access counts against both the
// vector-memory and the LDS counter, so every wait for an LDS access or a
// cross-lane shuffle (s_waitcnt lgkmcnt) also waits for every FLAT load in
// flight.  These helpers issue GLOBAL instructions, which only the
// vector-memory counter tracks.  The pointer must address global memory.
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
  if constexpr (sizeof(T) == 16) {
    typedef int v4 __attribute__((ext_vector_type(4)));
    return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) v4*)p);
  } else if constexpr (sizeof(T) == 8) {
    typedef int v2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) v2*)p);
  } else {
    static_assert(sizeof(T) == 4, "gld: 4, 8 or 16 bytes");
    return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) int*)p);
  }
}
// A pointer field of an argument struct, read as a global-address-space
// pointer: every access through the result (also inside inlined helpers that
// take plain pointers) compiles to GLOBAL instructions (see gld).  The field
// must hold a global-memory address or null.
template <typename T>
__device__ __forceinline__ T* GP(T* const& field) {
  __attribute__((address_space(1))) T* p;
  __builtin_memcpy(&p, &field, sizeof(p));
  return (T*)p;
}
template <typename T>
__device__ __forceinline__ void gst(T* p, const T& v) {
  if constexpr (sizeof(T) == 16) {
    typedef int v4 __attribute__((ext_vector_type(4)));
    *(__attribute__((address_space(1))) v4*)p = __builtin_bit_cast(v4, v);
  } else if constexpr (sizeof(T) == 8) {
    typedef int v2 __attribute__((ext_vector_type(2)));
    *(__attribute__((address_space(1))) v2*)p = __builtin_bit_cast(v2, v);
  } else {
    static_assert(sizeof(T) == 4, "gst: 4, 8 or 16 bytes");
    *(__attribute__((address_space(1))) int*)p = __builtin_bit_cast(int, v);
  }
}

// Cross-lane move of a double by DPP control CTRL (register-only, no LDS).
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffff), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double read_lane(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffff), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Sum over the 64 lanes of a wave, in a fixed order, returned to every lane.
// Quad permutes [1,0,3,2] and [2,3,0,1] give every lane its quad's sum, the
// half-row mirror (lane i of 8 reads lane 7 - i) the sum of its 8 lanes, the
// row mirror (lane i of 16 reads lane 15 - i) the sum of its row of 16; the
// four row sums are then read from lanes 0, 16, 32 and 48.  All lanes must be
// active.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_mov<0x141>(v);  // row_half_mirror
  v += dpp_mov<0x140>(v);  // row_mirror
  return ((read_lane(v, 0) + read_lane(v, 16)) + read_lane(v, 32)) + read_lane(v, 48);
}

// Fixed-order block reduction of NV doubles per thread; thread 0 of the block
// receives the sums.  blockDim.x must be a multiple of 64 and <= 1024, and
// every thread of the block must call it.
template <int NV>
__device__ inline void block_sum(double (&v)[NV], double* smem /* >= 16*NV */) {
#pragma unroll
  for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
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
