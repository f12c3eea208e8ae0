// mfista_kernels.hip -- fused elementwise MFISTA kernels and the FP64 control
// of the single-problem stream engine: one kernel per device function of
// mfista_dev.h (which cites the reference pieces they restate).
#include "mfista_dev.h"

namespace mf {

int elem_blocks(int n) {
  int b = (n + MF_BLOCK - 1) / MF_BLOCK;
  return b < 1 ? 1 : b;
}

template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_tsv_grad(const R* __restrict__ z, R* __restrict__ g, int W, int H, R lam, double* part,
           const int* skip) {
  if (skip && *skip) return;
  __shared__ double red[16];
  dev_tsv_grad<R>(blockIdx.x, z, g, W, H, lam, part, red);
}

template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_prox(const R* __restrict__ z, const R* __restrict__ g, R* __restrict__ xnew, R* __restrict__ dout,
       const float* __restrict__ box, int W, int H, const DevState* __restrict__ S, double* part,
       const int* skip) {
  if (skip && *skip) return;
  __shared__ double red[16 * kProxW];
  dev_prox<R>(blockIdx.x, z, g, xnew, dout, box, W, H, S, part, red);
}

template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_stats(const R* __restrict__ x, int W, int H, double* part) {
  __shared__ double red[64];
  dev_stats<R>(blockIdx.x, x, W, H, part, red);
}

template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_momentum(R* __restrict__ x, const R* __restrict__ xnew, R* __restrict__ z, int n,
           const DevState* __restrict__ S) {
  dev_momentum<R>(blockIdx.x, x, xnew, z, n, S);
}

template <typename Rs, typename Rd>
__global__ void k_cast(const Rs* __restrict__ a, Rd* __restrict__ b, int n, const int* skip) {
  if (skip && *skip) return;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n) b[f] = (Rd)a[f];
}

__global__ void __launch_bounds__(MF_BLOCK)
k_control(DevState* S, double* cost, const double* __restrict__ part, PartialLayout L) {
  __shared__ double red[16 * 9];
  dev_control(S, cost, part, L, red);
}

__global__ void __launch_bounds__(MF_BLOCK)
k_init_control(DevState* S, const double* __restrict__ part, PartialLayout L,
               const double* __restrict__ stats, int nsb) {
  __shared__ double red[16];
  dev_init_control(S, part, L, stats, nsb, red);
}

template <typename R>
void launch_tsv_grad(const R* z, R* g, int W, int H, double lambda_tsv, double* part,
                     hipStream_t st, const int* skip) {
  hipLaunchKernelGGL(k_tsv_grad<R>, dim3(elem_blocks(W * H)), dim3(MF_BLOCK), 0, st, z, g, W, H,
                     (R)lambda_tsv, part, skip);
}
template <typename R>
void launch_prox(const R* z, const R* g, R* xnew, R* d, const float* box, int W, int H,
                 const DevState* S, double* part, hipStream_t st, const int* skip) {
  hipLaunchKernelGGL(k_prox<R>, dim3(elem_blocks(W * H)), dim3(MF_BLOCK), 0, st, z, g, xnew, d, box,
                     W, H, S, part, skip);
}
template <typename R>
void launch_stats(const R* x, int W, int H, double* part, hipStream_t st) {
  hipLaunchKernelGGL(k_stats<R>, dim3(elem_blocks(W * H)), dim3(MF_BLOCK), 0, st, x, W, H, part);
}
template <typename R>
void launch_momentum(R* x, const R* xnew, R* z, int n, const DevState* S, hipStream_t st) {
  hipLaunchKernelGGL(k_momentum<R>, dim3(elem_blocks(n)), dim3(MF_BLOCK), 0, st, x, xnew, z, n, S);
}
void launch_control(DevState* S, double* cost, const double* part, PartialLayout L,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_control, dim3(1), dim3(MF_BLOCK), 0, st, S, cost, part, L);
}
void launch_init_control(DevState* S, const double* part, PartialLayout L, const double* stats,
                         int nsb, hipStream_t st) {
  hipLaunchKernelGGL(k_init_control, dim3(1), dim3(MF_BLOCK), 0, st, S, part, L, stats, nsb);
}
template <typename Rs, typename Rd>
void launch_cast(const Rs* a, Rd* b, int n, hipStream_t st, const int* skip) {
  hipLaunchKernelGGL((k_cast<Rs, Rd>), dim3(elem_blocks(n)), dim3(MF_BLOCK), 0, st, a, b, n, skip);
}

#define MF_INST2(R)                                                                          \
  template void launch_tsv_grad<R>(const R*, R*, int, int, double, double*, hipStream_t,     \
                                   const int*);                                              \
  template void launch_prox<R>(const R*, const R*, R*, R*, const float*, int, int,           \
                               const DevState*, double*, hipStream_t, const int*);           \
  template void launch_stats<R>(const R*, int, int, double*, hipStream_t);                   \
  template void launch_momentum<R>(R*, const R*, R*, int, const DevState*, hipStream_t);
MF_INST2(float)
MF_INST2(double)
template void launch_cast<float, double>(const float*, double*, int, hipStream_t, const int*);
template void launch_cast<double, float>(const double*, float*, int, hipStream_t, const int*);
template void launch_cast<double, double>(const double*, double*, int, hipStream_t, const int*);
template void launch_cast<float, float>(const float*, float*, int, hipStream_t, const int*);

}  // namespace mf
