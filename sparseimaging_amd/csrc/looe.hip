// looe.hip -- leave-one-out error estimate of a DFT-mode L1 or L1+TSV
// solution (c/looe_lib.c, Obuchi-Kabashima), FP64 on the GPU.
//
// With S the active set of x (|x_i| > 0) and A_s the columns of A in S:
//   H   = A_s^T A_s + lambda_tsv diag(d2_TSV(S))            (k_looe_hessian)
//   H   = L L^T                                             (k_chol_step, k_chol_scale)
//   d_m = a_m^T H^{-1} a_m = |L^{-1} a_m|^2                (k_looe_leverage)
//   s_m = ((y - A x)_m / (1 - d_m))^2 / 2
//   LOOE = mean(s), std = sqrt(sum s^2 / (M-1) - LOOE^2 M / (M-1))   (k_looe_sums + host)
// The reference solves H X = A_s^T with LAPACK dposv and forms d_m = a_m . X_m
// (c/looe_lib.c:87-168); one triangular solve gives the same d_m.  A pivot
// that is not positive marks H as not positive definite: LOOE = -1 with
// looe_m = looe_std = 0, as when dposv reports info > 0.
#include <cmath>
#include <vector>

#include "common.h"
#include "solver.h"

namespace mf {
namespace {

// d2_TSV (c/looe_lib.c:218-235): 8 inside, 6 on an edge, 4 at a corner, for
// pixel i at row i % NX, column i / NX; both indices are tested against NX,
// as the reference does.
__device__ inline double d2_tsv(int i, int NX) {
  const int r = i % NX, c = i / NX;
  const bool r_in = r > 0 && r < NX - 1, c_in = c > 0 && c < NX - 1;
  return r_in ? (c_in ? 8.0 : 6.0) : (c_in ? 6.0 : 4.0);
}

// H = A_s^T A_s (+ TSV diagonal), row-major n x n.  A_s is column-major M x n
// (column a at As + a M).  One 16 x 16 tile of H per 256-thread block; the M
// dimension goes through LDS in chunks of 16.
__global__ void __launch_bounds__(256) k_looe_hessian(const double* __restrict__ As, int M, int n,
                                                      const int* __restrict__ S, int NX, double lt,
                                                      double* __restrict__ H) {
  __shared__ double ta[16][17], tb[16][17];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int ra = blockIdx.y * 16 + ty, rb = blockIdx.x * 16 + ty;
  double acc = 0.0;
  for (int i0 = 0; i0 < M; i0 += 16) {
    const int i = i0 + tx;
    ta[ty][tx] = (ra < n && i < M) ? As[(size_t)ra * M + i] : 0.0;
    tb[ty][tx] = (rb < n && i < M) ? As[(size_t)rb * M + i] : 0.0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; k++) acc += ta[ty][k] * tb[tx][k];
    __syncthreads();
  }
  const int a = blockIdx.y * 16 + ty, b = blockIdx.x * 16 + tx;
  if (a < n && b < n) {
    if (a == b && lt > 0.0) acc += lt * d2_tsv(S[a], NX);
    H[(size_t)a * n + b] = acc;
  }
}

// Right-looking elimination step j of the Cholesky factorisation, lower
// triangle: H[i][k] -= H[i][j] H[k][j] / H[j][j] for j < k <= i < n.  Column j
// is read only, so the step has no races; after step j it holds L[.][j] times
// sqrt(H[j][j]).
__global__ void __launch_bounds__(256) k_chol_step(double* __restrict__ H, int n, int j, int* __restrict__ bad) {
  const double djj = H[(size_t)j * n + j];
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && !(djj > 0.0)) *bad = 1;
  const int k = j + 1 + blockIdx.x * 16 + (threadIdx.x & 15);
  const int i = j + 1 + blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i < n && k <= i) H[(size_t)i * n + k] -= H[(size_t)i * n + j] * H[(size_t)k * n + j] / djj;
}

// L[i][j] = H[i][j] / sqrt(H[j][j]) for j <= i (row-major, separate output).
__global__ void __launch_bounds__(256) k_chol_scale(const double* __restrict__ H, int n, double* __restrict__ L) {
  const int j = blockIdx.x * 16 + (threadIdx.x & 15);
  const int i = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i < n && j < n) L[(size_t)i * n + j] = (j <= i) ? H[(size_t)i * n + j] / sqrt(H[(size_t)j * n + j]) : 0.0;
}

// Thread m: y = L^{-1} a_m by forward substitution (a_m is row m of A_s; y
// overwrites it), d_m = |y|^2, s_m = ((y - A x)_m / (1 - d_m))^2 / 2.
__global__ void __launch_bounds__(64) k_looe_leverage(const double* __restrict__ L, int n, double* __restrict__ As,
                                                      int M, const double* __restrict__ yAx,
                                                      double* __restrict__ s) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  double d = 0.0;
  for (int i = 0; i < n; i++) {
    const double* Li = L + (size_t)i * n;
    double v = As[(size_t)i * M + m];
    for (int k = 0; k < i; k++) v -= Li[k] * As[(size_t)k * M + m];
    v /= Li[i];
    As[(size_t)i * M + m] = v;
    d += v * v;
  }
  const double t = yAx[m] / (1.0 - d);
  s[m] = t * t / 2.0;
}

// Fixed-order sums of s and s^2 (one block).
__global__ void __launch_bounds__(256) k_looe_sums(const double* __restrict__ s, int M, double* __restrict__ out) {
  __shared__ double red[16 * 2];
  double v[2] = {0.0, 0.0};
  for (int m = threadIdx.x; m < M; m += blockDim.x) {
    v[0] += s[m];
    v[1] += s[m] * s[m];
  }
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    out[0] = v[0];
    out[1] = v[1];
  }
}

struct LooeBufs {
  DevMem As, H, L, s, yAx, S, bad, sums;
  ~LooeBufs() {
    As.release();
    H.release();
    L.release();
    s.release();
    yAx.release();
    S.release();
    bad.release();
    sums.release();
  }
};

}  // namespace

// The LOOE of a DFT-mode solution x of y ~ A x (A: M x N column-major, host
// memory).  lambda_tsv = 0 is compute_LOOE_L1.  *pd = 1 if H is positive
// definite (then *looe_m, *looe_std are set), 0 if not (both 0).
int looe_dft(Context& C, int M, int N, int NX, double lambda_tsv, const double* A, const double* x,
             const double* yAx, double* looe_m, double* looe_std, int* pd, int* n_active) {
  if (M < 2 || N < 1 || !A || !x || !yAx || !looe_m || !looe_std || !pd) return MFISTA_E_ARG;
  if (lambda_tsv > 0.0 && NX < 1) return MFISTA_E_ARG;
  std::vector<int> S;  // find_active_set (c/looe_lib.c:52-66)
  for (int i = 0; i < N; i++)
    if (std::fabs(x[i]) > 0) S.push_back(i);
  const int n = (int)S.size();
  if (n_active) *n_active = n;
  // shrink_A (c/looe_lib.c:70-83) on the host: only the active columns travel
  std::vector<double> hAs((size_t)n * M);
  for (int j = 0; j < n; j++)
    std::copy(A + (size_t)S[j] * M, A + (size_t)S[j] * M + M, hAs.begin() + (size_t)j * M);
  LooeBufs b;
  const size_t nn = (size_t)(n > 0 ? n : 1) * (n > 0 ? n : 1);
  if (b.As.ensure((size_t)(n > 0 ? n : 1) * M * 8) || b.H.ensure(nn * 8) || b.L.ensure(nn * 8) ||
      b.s.ensure((size_t)M * 8) || b.yAx.ensure((size_t)M * 8) || b.S.ensure((size_t)(n > 0 ? n : 1) * 4) ||
      b.bad.ensure(4) || b.sums.ensure(16))
    return MFISTA_E_HIP;
  hipStream_t st = C.stream;
  if (n > 0) {
    MF_CHECK(hipMemcpyAsync(b.As.p, hAs.data(), hAs.size() * 8, hipMemcpyHostToDevice, st));
    MF_CHECK(hipMemcpyAsync(b.S.p, S.data(), (size_t)n * 4, hipMemcpyHostToDevice, st));
  }
  MF_CHECK(hipMemcpyAsync(b.yAx.p, yAx, (size_t)M * 8, hipMemcpyHostToDevice, st));
  MF_CHECK(hipMemsetAsync(b.bad.p, 0, 4, st));
  double* As = (double*)b.As.p;
  double* H = (double*)b.H.p;
  double* L = (double*)b.L.p;
  if (n > 0) {
    const int nt = (n + 15) / 16;
    hipLaunchKernelGGL(k_looe_hessian, dim3(nt, nt), dim3(256), 0, st, As, M, n, (const int*)b.S.p, NX,
                       lambda_tsv, H);
    for (int j = 0; j < n; j++) {
      const int t = (n - j - 1 + 15) / 16;
      hipLaunchKernelGGL(k_chol_step, dim3(t > 0 ? t : 1, t > 0 ? t : 1), dim3(256), 0, st, H, n, j,
                         (int*)b.bad.p);
    }
    hipLaunchKernelGGL(k_chol_scale, dim3(nt, nt), dim3(256), 0, st, (const double*)H, n, L);
  }
  hipLaunchKernelGGL(k_looe_leverage, dim3((M + 63) / 64), dim3(64), 0, st, (const double*)L, n, As, M,
                     (const double*)b.yAx.p, (double*)b.s.p);
  hipLaunchKernelGGL(k_looe_sums, dim3(1), dim3(256), 0, st, (const double*)b.s.p, M, (double*)b.sums.p);
  MF_CHECK(hipGetLastError());
  double sums[2];
  int bad = 0;
  MF_CHECK(hipMemcpyAsync(sums, b.sums.p, 16, hipMemcpyDeviceToHost, st));
  MF_CHECK(hipMemcpyAsync(&bad, b.bad.p, 4, hipMemcpyDeviceToHost, st));
  MF_CHECK(hipStreamSynchronize(st));
  if (bad) {  // c/looe_lib.c:160-166
    *looe_m = 0.0;
    *looe_std = 0.0;
    *pd = 0;
    return 0;
  }
  // c/looe_lib.c:147-156
  const double m = (double)M;
  const double lm = sums[0] / m;
  double var = sums[1] / (m - 1.0);
  var -= lm * lm * m / (m - 1.0);
  *looe_m = lm;
  *looe_std = std::sqrt(var);
  *pd = 1;
  return 0;
}

}  // namespace mf

using namespace mf;

extern "C" {

int mfista_looe_dft(int M, int N, int NX, int NY, double lambda_tsv, const double* Amat, const double* xvec,
                    const double* yAx, double* looe_m, double* looe_std, int* hessian_positive,
                    int* n_active) {
  (void)NY;
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  return looe_dft(*C, M, N, NX, lambda_tsv, Amat, xvec, yAx, looe_m, looe_std, hessian_positive, n_active);
}

static double looe_ref_api(int M, int N, int NX, int NY, double lambda_tsv, const double* Amat,
                           const double* xvec, const double* yAx, double* looe_m, double* looe_std) {
  int pd = 0;
  const int rc = mfista_looe_dft(M, N, NX, NY, lambda_tsv, Amat, xvec, yAx, looe_m, looe_std, &pd, nullptr);
  if (rc) {
    fprintf(stderr, "[mfista_amd] LOOE failed with status %d\n", rc);
    return -1.0;
  }
  return pd ? *looe_m : -1.0;
}

double compute_LOOE_L1(int* M, int* N, double lambda1, double* yvec, double* Amat, double* xvec, double* yAx,
                       double* looe_m, double* looe_std) {
  (void)lambda1;
  (void)yvec;
  return looe_ref_api(*M, *N, 1, 1, 0.0, Amat, xvec, yAx, looe_m, looe_std);
}

double compute_LOOE_L1_TSV(int* M, int* N, int NX, int NY, double lambda_l1, double lambda_tsv, double* yvec,
                           double* Amat, double* xvec, double* yAx, double* looe_m, double* looe_std) {
  (void)lambda_l1;
  (void)yvec;
  return looe_ref_api(*M, *N, NX, NY, lambda_tsv, Amat, xvec, yAx, looe_m, looe_std);
}

}  // extern "C"
