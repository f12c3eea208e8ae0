// dft_batch.hip -- DFT-mode MFISTA for many problems that share one matrix.
//
// The reference's DFT mode (c/mfista_dft_lib.c) multiplies a dense M x N
// matrix A (column-major, rows pre-weighted by 1/sigma) by one vector per
// product (dgemv, :30-48).  K problems that share A and y -- a grid of
// regularisation weights over one data set -- turn those products into GEMMs:
//   R = y 1^T - A X   (M x K)        G = A^T R   (N x K)
// which run here on the FP64 matrix cores (v_mfma_f64_16x16x4f64).  The K
// solves advance in lockstep rounds; each problem keeps its own step-size
// constant, momentum and backtracking, and takes exactly the decisions of the
// reference loop (c/mfista_dft_lib.c:282-436, L1 + TSV):
//   round = [A^T R_z for the problems that start an iteration] -> prox ->
//           [A X_new for every active problem] -> per-problem control.
// The residual of the momentum point is the same linear combination of the
// residuals of x and xnew as z is of x and xnew (the coefficients sum to one),
// so each iteration needs one adjoint and each trial one forward product.
//
// Layouts: A as the caller gives it (column-major, A[m + M n]); iterates and
// gradients N x Kp and residuals M x Kp with the problem index fastest
// (X[n Kp + k]), Kp = K rounded up to 16, 32 or 64, padding columns zero.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/mfista.h"
#include "solver.h"

namespace mf {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kDbMaxK = 64;   // problems per lockstep group (four 16-wide MFMA column tiles)
constexpr int kDbTileP = 64;  // output rows per workgroup (four waves of 16)
constexpr int kDbChunk = 16;  // inner-dimension chunk staged in LDS

// Split-inner FP64 GEMM on the matrix cores.  Workgroup (bx, s) computes rows
// [64 bx, 64 bx + 64) of the partial product over inner indices
// [s lchunk, (s + 1) lchunk) and writes part[s][p][k]:
//   MODE 0 (forward): P = M, L = N, op(A)(p, l) = A[p + M l]
//   MODE 1 (adjoint): P = N, L = M, op(A)(p, l) = A[l + M p]
//   B(l, k) = Bm[l Kp + k], k < Kp = 16 NKT.
// v_mfma_f64_16x16x4f64 lane layout (lane i of 64; measured on the MI355X
// with one-hot operands): A operand row i % 16, inner index i / 16; B operand
// inner index i / 16, column i % 16; result column i % 16, row i / 16 + 4 j in
// element j.
template <int MODE, int NKT>
__global__ void __launch_bounds__(256) k_dgemm_slices(const double* __restrict__ A, int M, int N,
                                                       const double* __restrict__ Bm, double* __restrict__ part,
                                                       int lchunk) {
  constexpr int KP = 16 * NKT;
  __shared__ double As[kDbChunk][kDbTileP + 1];
  __shared__ double Bs[kDbChunk][KP];
  const int P = MODE == 0 ? M : N, L = MODE == 0 ? N : M;
  const int p0 = blockIdx.x * kDbTileP, s = blockIdx.y;
  const int lbeg = s * lchunk, lend = min(L, lbeg + lchunk);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  v4d acc[NKT];
#pragma unroll
  for (int j = 0; j < NKT; j++) acc[j] = v4d{0.0, 0.0, 0.0, 0.0};
  for (int l0 = lbeg; l0 < lend; l0 += kDbChunk) {
#pragma unroll
    for (int j = 0; j < (kDbChunk * kDbTileP) / 256; j++) {
      const int idx = t + 256 * j;
      int pl, pp;
      if (MODE == 0) {  // 64 consecutive rows m per inner index n: coalesced along m
        pl = idx / kDbTileP;
        pp = idx - pl * kDbTileP;
      } else {  // 16 consecutive m per output row n
        pp = idx / kDbChunk;
        pl = idx - pp * kDbChunk;
      }
      const int p = p0 + pp, l = l0 + pl;
      double v = 0.0;
      if (p < P && l < lend) v = MODE == 0 ? A[(size_t)p + (size_t)M * l] : A[(size_t)l + (size_t)M * p];
      As[pl][pp] = v;
    }
    for (int idx = t; idx < kDbChunk * KP; idx += 256) {
      const int pl = idx / KP, kk = idx - pl * KP;
      const int l = l0 + pl;
      Bs[pl][kk] = (l < lend) ? Bm[(size_t)l * KP + kk] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int ls = 0; ls < kDbChunk; ls += 4) {
      const double a = As[ls + (lane >> 4)][w * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < NKT; j++) {
        const double b = Bs[ls + (lane >> 4)][j * 16 + (lane & 15)];
        acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < NKT; j++) {
#pragma unroll
    for (int e = 0; e < 4; e++) {
      const int p = p0 + w * 16 + (lane >> 4) + 4 * e;
      if (p < P) part[((size_t)s * P + p) * KP + j * 16 + (lane & 15)] = acc[j][e];
    }
  }
}

// Sums the slices in order.  Forward: R[m][k] = y[m] - sum_s part[s][m][k]
// and per-block partials of |R_k|^2 (psq[blk][k]); adjoint: G[n][k] = sum_s.
// Block = 256 threads over rows r0 .. r0 + 255 (thread: rows, problem k).
template <bool FWD>
__global__ void __launch_bounds__(256) k_dgemm_reduce(const double* __restrict__ part, int nslice, int P, int Kp,
                                                       const double* __restrict__ y, double* __restrict__ out,
                                                       double* __restrict__ psq) {
  __shared__ double red[256];
  const int t = threadIdx.x, k = t % Kp, rg = t / Kp, ng = 256 / Kp;
  const int r0 = blockIdx.x * 256;
  double sq = 0.0;
  for (int r = r0 + rg; r < min(P, r0 + 256); r += ng) {
    double v = 0.0;
    for (int s = 0; s < nslice; s++) v += part[((size_t)s * P + r) * Kp + k];
    if (FWD) {
      v = y[r] - v;
      sq += v * v;
    }
    out[(size_t)r * Kp + k] = v;
  }
  if (FWD) {
    red[t] = sq;
    __syncthreads();
    if (rg == 0) {
      double a = 0.0;
      for (int q = 0; q < ng; q++) a += red[q * Kp + k];
      psq[(size_t)blockIdx.x * Kp + k] = a;
    }
  }
}

// Per-problem state of the lockstep solver (reference loop variables).
struct DbState {
  double c, mu, costtmp, l1x, Qcore, ma, mb;
  double lambda_l1, lambda_tsv, eps, rel_tol;
  int iter, trial, need_grad, active, accepted, finished, ITER, nonneg, maxiter, max_trials, total_trials;
  double ring[kTD + 1];
};

// partial-sum slots per element block (per problem): TSV(z), |xnew|_1,
// <d, dfdx>, |d|^2, TSV(xnew)
enum { kDbTsvZ = 0, kDbL1 = 1, kDbDg = 2, kDbD2 = 3, kDbTsvX = 4, kDbNQ = 5 };

// TSV term and its gradient at flat pixel f of the x[NX j + i] view
// (c/mfista_tools.c:145-186): the term counts the right and lower
// neighbours, the gradient 2 sum over all four of (x_f - x_n).
__device__ inline void db_tsv_at(const double* X, int Kp, int k, int f, int NX, int NY, double& term, double& grad) {
  const int i = f % NX, j = f / NX;
  const double xc = X[(size_t)f * Kp + k];
  term = 0.0;
  grad = 0.0;
  if (i < NX - 1) {
    const double d = xc - X[(size_t)(f + 1) * Kp + k];
    term += d * d;
    grad += 2.0 * d;
  }
  if (i > 0) grad += 2.0 * (xc - X[(size_t)(f - 1) * Kp + k]);
  if (j < NY - 1) {
    const double d = xc - X[(size_t)(f + NX) * Kp + k];
    term += d * d;
    grad += 2.0 * d;
  }
  if (j > 0) grad += 2.0 * (xc - X[(size_t)(f - NX) * Kp + k]);
}

// Block-level fixed-order sum of NQ values per problem: thread (rg, k) holds
// its values; psum[q][blk][k].
template <int NQ>
__device__ inline void db_block_sums(double (&v)[NQ], int Kp, double* __restrict__ psum, int nblk, double* red) {
  const int t = threadIdx.x, k = t % Kp, rg = t / Kp, ng = 256 / Kp;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    red[t] = v[q];
    __syncthreads();
    if (rg == 0) {
      double a = 0.0;
      for (int g = 0; g < ng; g++) a += red[g * Kp + k];
      psum[((size_t)q * nblk + blockIdx.x) * Kp + k] = a;
    }
    __syncthreads();
  }
}

// Gradient (for problems that start an iteration) and prox: dfdx = A^T R_z -
// lambda_tsv dTSV(z); xnew = prox(z + dfdx / c); partials of TSV(z), |xnew|_1,
// <xnew - z, dfdx>, |xnew - z|^2.  Blocks over pixels, threads (pixel, k).
__global__ void __launch_bounds__(256) k_db_prox(const double* __restrict__ G, const double* __restrict__ Z,
                                                  double* __restrict__ D, double* __restrict__ XN,
                                                  const DbState* __restrict__ S, const float* __restrict__ box,
                                                  int N, int NX, int NY, int Kp, int K, double* __restrict__ psum,
                                                  int nblk) {
  __shared__ double red[256];
  const int t = threadIdx.x, k = t % Kp, rg = t / Kp, ng = 256 / Kp;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const bool live = k < K && S[k].active;
  const int f0 = blockIdx.x * 256;
  if (live) {
    const DbState& s = S[k];
    const double cinv = 1.0 / s.c, eta = s.lambda_l1 / s.c;
    for (int f = f0 + rg; f < min(N, f0 + 256); f += ng) {
      const size_t e = (size_t)f * Kp + k;
      double g;
      if (s.need_grad) {
        double term, grad;
        db_tsv_at(Z, Kp, k, f, NX, NY, term, grad);
        g = G[e];
        if (s.lambda_tsv > 0.0) {
          g -= s.lambda_tsv * grad;
          v[kDbTsvZ] += term;
        }
        D[e] = g;
      } else {
        g = D[e];
      }
      const double zf = Z[e], tmp = g * cinv + zf;
      double xn;
      if (box && box[f] == 0.0f)
        xn = 0.0;
      else if (s.nonneg == 1)
        xn = tmp > eta ? tmp - eta : 0.0;
      else
        xn = tmp > eta ? tmp - eta : (tmp < -eta ? tmp + eta : 0.0);
      XN[e] = xn;
      const double d = xn - zf;
      v[kDbL1] += fabs(xn);
      v[kDbDg] += d * g;
      v[kDbD2] += d * d;
    }
  }
  db_block_sums<4>(v, Kp, psum, nblk, red);
}

// TSV(xnew) partials (slot kDbTsvX)
__global__ void __launch_bounds__(256) k_db_tsv(const double* __restrict__ XN, const DbState* __restrict__ S, int N,
                                                 int NX, int NY, int Kp, int K, double* __restrict__ psum, int nblk) {
  __shared__ double red[256];
  const int t = threadIdx.x, k = t % Kp, rg = t / Kp, ng = 256 / Kp;
  double v[1] = {0.0};
  if (k < K && S[k].active && S[k].lambda_tsv > 0.0) {
    const int f0 = blockIdx.x * 256;
    for (int f = f0 + rg; f < min(N, f0 + 256); f += ng) {
      double term, grad;
      db_tsv_at(XN, Kp, k, f, NX, NY, term, grad);
      v[0] += term;
    }
  }
  db_block_sums<1>(v, Kp, psum + (size_t)kDbTsvX * nblk * Kp, nblk, red);
}

// Control (one thread per problem): the decisions of the reference loop,
// c/mfista_dft_lib.c:282-436, from fixed-order sums of the block partials.
__global__ void k_db_control(DbState* __restrict__ S, int K, const double* __restrict__ psum, int nblk_n,
                             const double* __restrict__ psq_n, const double* __restrict__ psq_z, int nblk_m,
                             int Kp, double* __restrict__ cost_hist, int maxiter_all) {
  const int k = threadIdx.x;
  if (k >= K) return;
  DbState& s = S[k];
  s.finished = 0;  // set again below when this round ends an iteration
  s.accepted = 0;
  if (!s.active) return;
  auto sum_n = [&](int q) {
    double a = 0.0;
    for (int b = 0; b < nblk_n; b++) a += psum[((size_t)q * nblk_n + b) * Kp + k];
    return a;
  };
  double chi_n = 0.0, chi_z = 0.0;
  for (int b = 0; b < nblk_m; b++) chi_n += psq_n[(size_t)b * Kp + k];
  if (s.need_grad) {  // a new iteration: Qcore = F(z) + lambda_tsv TSV(z)
    for (int b = 0; b < nblk_m; b++) chi_z += psq_z[(size_t)b * Kp + k];
    s.Qcore = chi_z / 2 + (s.lambda_tsv > 0.0 ? s.lambda_tsv * sum_n(kDbTsvZ) : 0.0);
    s.ring[s.iter % (kTD + 1)] = s.costtmp;
    cost_hist[(size_t)k * maxiter_all + s.iter] = s.costtmp;
  }
  const double l1n = sum_n(kDbL1), dg = sum_n(kDbDg), d2 = sum_n(kDbD2);
  double F = chi_n / 2 + (s.lambda_tsv > 0.0 ? s.lambda_tsv * sum_n(kDbTsvX) : 0.0);
  const double Q = -dg + s.c * d2 / 2 + s.Qcore;
  s.trial += 1;
  s.total_trials += 1;
  if (!(F <= Q) && s.trial < s.maxiter) {  // c/mfista_dft_lib.c: c *= ETA, next trial
    s.c *= kEta;
    s.need_grad = 0;
    return;
  }
  if (!(F <= Q)) s.c *= kEta;
  if (s.trial > s.max_trials) s.max_trials = s.trial;
  s.trial = 0;
  s.c /= kEta;
  const double mu = s.mu, munew = (1 + sqrt(1 + 4 * mu * mu)) / 2;
  F += s.lambda_l1 * l1n;
  const int it = s.iter;
  const double cit = s.ring[it % (kTD + 1)];
  bool stop = false;
  if (F < cit) {
    s.costtmp = F;
    s.ma = (1 - mu) / munew;
    s.mb = 1 + (mu - 1) / munew;
    s.accepted = 1;
    s.l1x = l1n;
  } else {
    s.ma = 1 - mu / munew;
    s.mb = mu / munew;
    if (it > 1 && s.l1x == 0.0) stop = true;
  }
  if (!stop && it >= kMinIter) {
    const double d = s.ring[(it - kTD) % (kTD + 1)] - cit;
    stop = (s.rel_tol > 0) ? (d <= s.rel_tol * cit) : (d < s.eps);
  }
  s.mu = munew;
  s.finished = 1;
  s.need_grad = 1;
  if (stop) {
    s.active = 0;
    s.ITER = it + 1;
  } else if (it + 1 >= s.maxiter) {
    s.active = 0;
    s.ITER = s.maxiter;
  } else {
    s.iter = it + 1;
  }
}

// The end of an iteration (finished problems): z = a x + b xnew, x = xnew on
// acceptance, and the same for the residuals; |R_z|^2 partials for the next
// iteration's Qcore.  rows: N pixels (images) then M visibilities (residuals).
__global__ void __launch_bounds__(256) k_db_update(double* __restrict__ X, double* __restrict__ Z,
                                                    const double* __restrict__ XN, double* __restrict__ Rx,
                                                    double* __restrict__ Rz, const double* __restrict__ Rn,
                                                    const DbState* __restrict__ S, int N, int M, int Kp, int K,
                                                    int nblk_img, double* __restrict__ psq_z) {
  __shared__ double red[256];
  const int t = threadIdx.x, k = t % Kp, rg = t / Kp, ng = 256 / Kp;
  const bool fin = k < K && S[k].finished;
  const double a = fin ? S[k].ma : 0.0, b = fin ? S[k].mb : 0.0;
  const bool acc = fin && S[k].accepted;
  if ((int)blockIdx.x < nblk_img) {
    const int f0 = blockIdx.x * 256;
    if (fin)
      for (int f = f0 + rg; f < min(N, f0 + 256); f += ng) {
        const size_t e = (size_t)f * Kp + k;
        const double xf = X[e], xn = XN[e];
        Z[e] = a * xf + b * xn;
        if (acc) X[e] = xn;
      }
    return;
  }
  const int blk = blockIdx.x - nblk_img, m0 = blk * 256;
  double sq = 0.0;
  if (fin)
    for (int m = m0 + rg; m < min(M, m0 + 256); m += ng) {
      const size_t e = (size_t)m * Kp + k;
      const double rx = Rx[e], rn = Rn[e];
      const double rz = a * rx + b * rn;
      Rz[e] = rz;
      sq += rz * rz;
      if (acc) Rx[e] = rn;
    }
  red[t] = sq;
  __syncthreads();
  if (rg == 0) {
    double s = 0.0;
    for (int g = 0; g < ng; g++) s += red[g * Kp + k];
    psq_z[(size_t)blk * Kp + k] = s;
  }
}

// ---------------------------------------------------------------------------
struct DbWork {
  int M = 0, N = 0, Kp = 0, nsl_f = 0, nsl_a = 0, lch_f = 0, lch_a = 0;
};

static DbWork db_plan(int M, int N, int K) {
  DbWork w;
  w.M = M;
  w.N = N;
  w.Kp = 16;  // a power of two: 256 / Kp row groups of Kp threads tile a block exactly
  while (w.Kp < K) w.Kp *= 2;
  int dev = 0, ncu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
  // slices of the inner dimension so that a product has about 2 workgroups per CU
  auto slices = [&](int P, int L, int& nsl, int& lch) {
    const int tiles = (P + kDbTileP - 1) / kDbTileP;
    nsl = std::max(1, std::min((2 * ncu + tiles - 1) / tiles, (L + 4 * kDbChunk - 1) / (4 * kDbChunk)));
    lch = ((L + nsl - 1) / nsl + kDbChunk - 1) / kDbChunk * kDbChunk;
    nsl = (L + lch - 1) / lch;
  };
  slices(M, N, w.nsl_f, w.lch_f);
  slices(N, M, w.nsl_a, w.lch_a);
  return w;
}

template <int MODE>
static void launch_gemm(const DbWork& w, const double* A, const double* B, double* part, hipStream_t st) {
  const int P = MODE == 0 ? w.M : w.N;
  const dim3 grid((P + kDbTileP - 1) / kDbTileP, MODE == 0 ? w.nsl_f : w.nsl_a);
  const int lch = MODE == 0 ? w.lch_f : w.lch_a;
  switch (w.Kp / 16) {
    case 1: hipLaunchKernelGGL((k_dgemm_slices<MODE, 1>), grid, dim3(256), 0, st, A, w.M, w.N, B, part, lch); break;
    case 2: hipLaunchKernelGGL((k_dgemm_slices<MODE, 2>), grid, dim3(256), 0, st, A, w.M, w.N, B, part, lch); break;
    default: hipLaunchKernelGGL((k_dgemm_slices<MODE, 4>), grid, dim3(256), 0, st, A, w.M, w.N, B, part, lch); break;
  }
}

// R = y 1^T - A X (psq: |R_k|^2 partials per 256-row block)
static void db_forward(const DbWork& w, const double* A, const double* X, const double* y, double* part, double* R,
                       double* psq, hipStream_t st) {
  launch_gemm<0>(w, A, X, part, st);
  hipLaunchKernelGGL(k_dgemm_reduce<true>, dim3((w.M + 255) / 256), dim3(256), 0, st, (const double*)part, w.nsl_f,
                     w.M, w.Kp, y, R, psq);
}

// G = A^T R
static void db_adjoint(const DbWork& w, const double* A, const double* R, double* part, double* G, hipStream_t st) {
  launch_gemm<1>(w, A, R, part, st);
  hipLaunchKernelGGL(k_dgemm_reduce<false>, dim3((w.N + 255) / 256), dim3(256), 0, st, (const double*)part,
                     w.nsl_a, w.N, w.Kp, (const double*)nullptr, G, (double*)nullptr);
}

}  // namespace mf

using namespace mf;

extern "C" int mfista_dft_products(const double* A, int M, int N, int K, const double* X, const double* y, double* R,
                                   double* G) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (!A || !X || !R || M <= 0 || N <= 0 || K <= 0 || K > kDbMaxK) return MFISTA_E_ARG;
  const DbWork w = db_plan(M, N, K);
  hipStream_t st = C->stream;
  DftBatchBufs& B = C->dbb;
  const size_t kp = (size_t)w.Kp;
  std::vector<double> xt((size_t)N * kp, 0.0), yv((size_t)M, 0.0);
  for (int k = 0; k < K; k++)
    for (int n = 0; n < N; n++) xt[(size_t)n * kp + k] = X[(size_t)k * N + n];
  if (y) std::copy(y, y + M, yv.begin());
  const size_t part_el = std::max((size_t)w.nsl_f * M, (size_t)w.nsl_a * N) * kp;
  int rc;
  if ((rc = upload(*C, B.A, A, (size_t)M * N * 8)) || (rc = upload(*C, B.X, xt.data(), xt.size() * 8)) ||
      (rc = upload(*C, B.y, yv.data(), (size_t)M * 8)))
    return rc;
  if (B.Rx.ensure((size_t)M * kp * 8) || B.G.ensure((size_t)N * kp * 8) || B.part.ensure(part_el * 8) ||
      B.psq_n.ensure((size_t)((M + 255) / 256) * kp * 8))
    return MFISTA_E_HIP;
  db_forward(w, (const double*)B.A.p, (const double*)B.X.p, (const double*)B.y.p, (double*)B.part.p,
             (double*)B.Rx.p, (double*)B.psq_n.p, st);
  if (G) db_adjoint(w, (const double*)B.A.p, (const double*)B.Rx.p, (double*)B.part.p, (double*)B.G.p, st);
  std::vector<double> rt((size_t)M * kp), gt(G ? (size_t)N * kp : 0);
  MF_CHECK(hipMemcpyAsync(rt.data(), B.Rx.p, rt.size() * 8, hipMemcpyDeviceToHost, st));
  if (G) MF_CHECK(hipMemcpyAsync(gt.data(), B.G.p, gt.size() * 8, hipMemcpyDeviceToHost, st));
  MF_CHECK(hipStreamSynchronize(st));
  MF_CHECK(hipGetLastError());
  for (int k = 0; k < K; k++) {
    for (int m = 0; m < M; m++) R[(size_t)k * M + m] = rt[(size_t)m * kp + k];
    if (G)
      for (int n = 0; n < N; n++) G[(size_t)k * N + n] = gt[(size_t)n * kp + k];
  }
  return 0;
}

extern "C" int mfista_imaging_core_dft_batch(const double* y, const double* A, int M, int N, int NX, int NY,
                                             int nprob, const double* lambda_l1, const double* lambda_tsv,
                                             const double* cinit, int maxiter, double eps, const double* xinit,
                                             double* xout, int nonneg_flag, int box_flag, const float* cl_box,
                                             struct RESULT* results, mfista_opts* opts) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (!y || !A || !lambda_l1 || !lambda_tsv || !cinit || !xout || !results || nprob <= 0 || M <= 0 || N <= 0 ||
      N != NX * NY || maxiter <= 0 || (box_flag && !cl_box))
    return MFISTA_E_ARG;
  if (nonneg_flag != 0 && nonneg_flag != 1) {  // c/mfista_dft_lib.c: the core returns 0 iterations
    printf("nonneg_flag must be chosen properly.\n");
    for (int k = 0; k < nprob; k++) {
      memset(&results[k], 0, sizeof(RESULT));
      results[k].maxiter = maxiter;
      results[k].nonneg = nonneg_flag;
    }
    return MFISTA_E_NONNEG;
  }
  hipStream_t st = C->stream;
  const double rel_tol = opts ? opts->rel_tol : 0.0;
  auto t0 = std::chrono::steady_clock::now();
  DftBatchBufs& B = C->dbb;
  int rc;
  if ((rc = upload(*C, B.A, A, (size_t)M * N * 8)) || (rc = upload(*C, B.y, y, (size_t)M * 8))) return rc;
  if (box_flag && (rc = upload(*C, B.box, cl_box, (size_t)N * 4))) return rc;
  int total_trials = 0, max_trials = 0;
  // lockstep groups of up to kDbMaxK problems
  for (int g0 = 0; g0 < nprob; g0 += kDbMaxK) {
    const int K = std::min(kDbMaxK, nprob - g0);
    const DbWork w = db_plan(M, N, K);
    const size_t kp = (size_t)w.Kp;
    const int nblk_n = (N + 255) / 256, nblk_m = (M + 255) / 256;
    const size_t part_el = std::max((size_t)w.nsl_f * M, (size_t)w.nsl_a * N) * kp;
    if (B.X.ensure((size_t)N * kp * 8) || B.Z.ensure((size_t)N * kp * 8) || B.XN.ensure((size_t)N * kp * 8) ||
        B.D.ensure((size_t)N * kp * 8) || B.G.ensure((size_t)N * kp * 8) || B.Rx.ensure((size_t)M * kp * 8) ||
        B.Rz.ensure((size_t)M * kp * 8) || B.Rn.ensure((size_t)M * kp * 8) || B.part.ensure(part_el * 8) ||
        B.psum.ensure((size_t)kDbNQ * nblk_n * kp * 8) || B.psq_n.ensure((size_t)nblk_m * kp * 8) ||
        B.psq_z.ensure((size_t)nblk_m * kp * 8) || B.S.ensure(sizeof(DbState) * kDbMaxK) ||
        B.cost.ensure((size_t)K * maxiter * 8))
      return MFISTA_E_HIP;
    // x0 (problem-major from the caller) -> N x Kp
    std::vector<double> xt((size_t)N * kp, 0.0);
    for (int k = 0; k < K; k++)
      for (int n = 0; n < N; n++) xt[(size_t)n * kp + k] = xinit ? xinit[(size_t)(g0 + k) * N + n] : 0.0;
    MF_CHECK(hipMemcpyAsync(B.X.p, xt.data(), xt.size() * 8, hipMemcpyHostToDevice, st));
    MF_CHECK(hipMemcpyAsync(B.Z.p, xt.data(), xt.size() * 8, hipMemcpyHostToDevice, st));
    double* X = (double*)B.X.p;
    double* Z = (double*)B.Z.p;
    double* XN = (double*)B.XN.p;
    double* D = (double*)B.D.p;
    double* G = (double*)B.G.p;
    double* Rx = (double*)B.Rx.p;
    double* Rz = (double*)B.Rz.p;
    double* Rn = (double*)B.Rn.p;
    double* part = (double*)B.part.p;
    double* psum = (double*)B.psum.p;
    double* psq_n = (double*)B.psq_n.p;
    double* psq_z = (double*)B.psq_z.p;
    const double* dA = (const double*)B.A.p;
    const double* dy = (const double*)B.y.p;
    const float* dbox = box_flag ? (const float*)B.box.p : nullptr;
    // initial residual and cost: R_x = y - A x0, cost0 = F + lambda_l1 |x|_1 + lambda_tsv TSV(x)
    db_forward(w, dA, X, dy, part, Rx, psq_n, st);
    MF_CHECK(hipMemcpyAsync(Rz, Rx, (size_t)M * kp * 8, hipMemcpyDeviceToDevice, st));
    MF_CHECK(hipMemcpyAsync(psq_z, psq_n, (size_t)nblk_m * kp * 8, hipMemcpyDeviceToDevice, st));
    std::vector<double> hq((size_t)nblk_m * kp);
    MF_CHECK(hipMemcpyAsync(hq.data(), psq_n, hq.size() * 8, hipMemcpyDeviceToHost, st));
    MF_CHECK(hipStreamSynchronize(st));
    std::vector<DbState> hs(K);
    for (int k = 0; k < K; k++) {
      const double* x0 = xinit ? xinit + (size_t)(g0 + k) * N : nullptr;
      double chi = 0.0, l1 = 0.0, tsv = 0.0;
      for (int b = 0; b < nblk_m; b++) chi += hq[(size_t)b * kp + k];
      if (x0) {
        for (int n = 0; n < N; n++) l1 += fabs(x0[n]);
        for (int j = 0; j < NY; j++)  // TSV in the x[NX j + i] view, c/mfista_tools.c:145-163
          for (int i = 0; i < NX; i++) {
            const double xc = x0[NX * j + i];
            if (i < NX - 1) tsv += (xc - x0[NX * j + i + 1]) * (xc - x0[NX * j + i + 1]);
            if (j < NY - 1) tsv += (xc - x0[NX * (j + 1) + i]) * (xc - x0[NX * (j + 1) + i]);
          }
      }
      DbState& s = hs[k];
      memset(&s, 0, sizeof(s));
      s.c = cinit[g0 + k];
      s.mu = 1.0;
      s.lambda_l1 = lambda_l1[g0 + k];
      s.lambda_tsv = lambda_tsv[g0 + k];
      s.costtmp = chi / 2 + s.lambda_l1 * l1 + (s.lambda_tsv > 0 ? s.lambda_tsv * tsv : 0.0);
      s.l1x = l1;
      s.eps = eps;  // unscaled, c/mfista_dft_lib.c:559
      s.rel_tol = rel_tol;
      s.nonneg = nonneg_flag;
      s.maxiter = maxiter;
      s.need_grad = 1;
      s.active = 1;
    }
    MF_CHECK(hipMemcpyAsync(B.S.p, hs.data(), sizeof(DbState) * K, hipMemcpyHostToDevice, st));
    DbState* dS = (DbState*)B.S.p;
    // lockstep rounds; the host checks for completion every few rounds
    const long max_rounds = (long)maxiter * 64 + 1000;
    long round = 0;
    std::vector<DbState> chk(K);
    while (true) {
      for (int r = 0; r < 8; r++, round++) {
        db_adjoint(w, dA, Rz, part, G, st);
        hipLaunchKernelGGL(k_db_prox, dim3(nblk_n), dim3(256), 0, st, (const double*)G, (const double*)Z, D, XN,
                           (const DbState*)dS, dbox, N, NX, NY, (int)kp, K, psum, nblk_n);
        db_forward(w, dA, XN, dy, part, Rn, psq_n, st);
        hipLaunchKernelGGL(k_db_tsv, dim3(nblk_n), dim3(256), 0, st, (const double*)XN, (const DbState*)dS, N, NX,
                           NY, (int)kp, K, psum, nblk_n);
        hipLaunchKernelGGL(k_db_control, dim3(1), dim3(kDbMaxK), 0, st, dS, K, (const double*)psum, nblk_n,
                           (const double*)psq_n, (const double*)psq_z, nblk_m, (int)kp, (double*)B.cost.p, maxiter);
        hipLaunchKernelGGL(k_db_update, dim3(nblk_n + nblk_m), dim3(256), 0, st, X, Z, (const double*)XN, Rx, Rz,
                           (const double*)Rn, (const DbState*)dS, N, M, (int)kp, K, nblk_n, psq_z);
      }
      MF_CHECK(hipMemcpyAsync(chk.data(), dS, sizeof(DbState) * K, hipMemcpyDeviceToHost, st));
      MF_CHECK(hipStreamSynchronize(st));
      MF_CHECK(hipGetLastError());
      bool any = false;
      for (const DbState& s : chk) any = any || s.active;
      if (!any) break;
      if (round > max_rounds) return MFISTA_E_HIP;  // cannot happen for a convergent run
    }
    // note: a round after all problems finished is a no-op (inactive problems are skipped)
    std::vector<double> xo((size_t)N * kp), ro((size_t)M * kp);
    MF_CHECK(hipMemcpyAsync(xo.data(), X, xo.size() * 8, hipMemcpyDeviceToHost, st));
    MF_CHECK(hipMemcpyAsync(ro.data(), Rx, ro.size() * 8, hipMemcpyDeviceToHost, st));
    MF_CHECK(hipStreamSynchronize(st));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    for (int k = 0; k < K; k++) {
      double* xk = xout + (size_t)(g0 + k) * N;
      for (int n = 0; n < N; n++) xk[n] = xo[(size_t)n * kp + k];
      // RESULT as calc_result (c/mfista_dft_lib.c:440-536), FP64 on the host
      RESULT& R = results[g0 + k];
      memset(&R, 0, sizeof(R));
      const DbState& s = chk[k];
      double sq = 0.0, l1 = 0.0, tsv = 0.0;
      int nact = 0;
      R.residual = (double*)malloc(sizeof(double) * (size_t)M);
      if (!R.residual) return MFISTA_E_ARG;
      for (int m = 0; m < M; m++) {
        const double r = ro[(size_t)m * kp + k];
        R.residual[m] = r;
        sq += r * r;
      }
      for (int n = 0; n < N; n++) {
        l1 += fabs(xk[n]);
        if (xk[n] != 0.0) nact++;
      }
      for (int j = 0; j < NY; j++)
        for (int i = 0; i < NX; i++) {
          const double xc = xk[NX * j + i];
          if (i < NX - 1) tsv += (xc - xk[NX * j + i + 1]) * (xc - xk[NX * j + i + 1]);
          if (j < NY - 1) tsv += (xc - xk[NX * (j + 1) + i]) * (xc - xk[NX * (j + 1) + i]);
        }
      R.M = M;
      R.N = N;
      R.NX = NX;
      R.NY = NY;
      R.N_active = nact;
      R.maxiter = maxiter;
      R.ITER = s.ITER;
      R.nonneg = nonneg_flag;
      R.lambda_l1 = s.lambda_l1;
      R.lambda_tsv = s.lambda_tsv;
      R.sq_error = sq;
      R.mean_sq_error = sq / (double)M;
      R.l1cost = l1;
      R.finalcost = sq / 2 + s.lambda_l1 * l1;
      if (s.lambda_tsv > 0) {
        R.tsvcost = tsv;
        R.finalcost += s.lambda_tsv * tsv;
      }
      R.comp_time = secs;
      R.Lip_const = s.c;
      total_trials += s.total_trials;
      max_trials = std::max(max_trials, s.max_trials);
    }
  }
  if (opts) {
    opts->solve_ms = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() * 1e3;
    opts->total_trials = total_trials;
    opts->max_trials = max_trials;
  }
  return 0;
}
