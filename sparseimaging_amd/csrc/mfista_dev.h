// mfista_dev.h -- the MFISTA outer-loop computations as device functions:
// fused elementwise passes (TSV gradient, prox, image statistics, momentum) and
// the FP64 control step.  Each processes one block of 256 elements (`blk`) or,
// for the control, a whole problem with one workgroup; `red` is the caller's
// LDS scratch for the block reductions.  The single-problem kernels
// (mfista_kernels.hip) and the batched kernels of the wide engine (wide.hip)
// both call them.
//
// Reference pieces restated here:
//   TSV, d_TSV              c/mfista_tools.c:145-186 (x[NX*j+i] view: W = NX
//                           columns, H = NY rows of the flat image)
//   soft_threshold*_box     c/mfista_tools.c:81-141
//   calc_Q_part             c/mfista_tools.c:61-77
//   loop, backtracking, monotone momentum, stopping rules
//                           c/mfista_nufft_lib.c:660-752
#pragma once

#include "mfista_kernels.h"

namespace mf {

// g -= lambda_tsv * dTSV(z); per-block partial of TSV(z)
template <typename R>
__device__ inline void dev_tsv_grad(int blk, const R* __restrict__ z, R* __restrict__ g, int W, int H, R lam,
                                    double* part, double* red) {
  const int n = W * H;
  const int f = blk * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (f < n) {
    const int c = f % W, r = f / W;
    const R zc = z[f];
    R d = 0;
    if (c < W - 1) {
      const R t = zc - z[f + 1];
      d += 2 * t;
      v[0] += (double)t * (double)t;
    }
    if (c > 0) d += 2 * (zc - z[f - 1]);
    if (r < H - 1) {
      const R t = zc - z[f + W];
      d += 2 * t;
      v[0] += (double)t * (double)t;
    }
    if (r > 0) d += 2 * (zc - z[f - W]);
    g[f] = g[f] + (-lam) * d;
  }
  block_sum<1>(v, red);
  if (threadIdx.x == 0) part[blk] = v[0];
}

template <typename R>
__device__ inline R prox1(R v, R eta, int nonneg, bool boxed_out) {
  if (boxed_out) return 0;
  if (nonneg) return (v < eta) ? (R)0 : v - eta;
  if (v >= eta) return v - eta;
  if (v <= -eta) return v + eta;
  return 0;
}

// xnew = prox(z + g/c), d = xnew - z; partials: <d, g>, |d|^2, |xnew|_1,
// TSV(xnew), TSV(d)
template <typename R>
__device__ inline void dev_prox(int blk, const R* __restrict__ z, const R* __restrict__ g, R* __restrict__ xnew,
                                R* __restrict__ dout, const float* __restrict__ box, int W, int H,
                                const DevState* __restrict__ S, double* part, double* red) {
  const int n = W * H;
  const int f = blk * blockDim.x + threadIdx.x;
  const double c = S->c;
  const R cinv = (R)(1.0 / c);
  const R eta = (R)(S->lambda_l1 / c);
  const int nonneg = S->nonneg;
  const bool tsv = S->lambda_tsv > 0.0;
  double v[kProxW] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (f < n) {
    const R gf = g[f], zf = z[f];
    const R xn = prox1<R>(gf * cinv + zf, eta, nonneg, box && box[f] == 0.0f);
    xnew[f] = xn;
    const R d = xn - zf;
    if (dout) dout[f] = d;
    v[0] = (double)gf * (double)d;
    v[1] = (double)d * (double)d;
    v[2] = fabs((double)xn);
    if (tsv) {
      const int cc = f % W, r = f / W;
      if (cc < W - 1) {
        const R xr = prox1<R>(g[f + 1] * cinv + z[f + 1], eta, nonneg, box && box[f + 1] == 0.0f);
        const double t = (double)xn - (double)xr;
        const double u = (double)d - ((double)xr - (double)z[f + 1]);
        v[3] += t * t;
        v[4] += u * u;
      }
      if (r < H - 1) {
        const R xd = prox1<R>(g[f + W] * cinv + z[f + W], eta, nonneg, box && box[f + W] == 0.0f);
        const double t = (double)xn - (double)xd;
        const double u = (double)d - ((double)xd - (double)z[f + W]);
        v[3] += t * t;
        v[4] += u * u;
      }
    }
  }
  block_sum<kProxW>(v, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < kProxW; q++) part[kProxW * blk + q] = v[q];
}

// |x|_1, TSV(x), TV(x), #(x != 0) partials (FP64 arithmetic on R inputs)
template <typename R>
__device__ inline void dev_stats(int blk, const R* __restrict__ x, int W, int H, double* part, double* red) {
  const int n = W * H;
  const int f = blk * blockDim.x + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (f < n) {
    const int c = f % W, r = f / W;
    const double xc = (double)x[f];
    v[0] = fabs(xc);
    v[3] = (xc != 0.0) ? 1.0 : 0.0;
    double dh = 0, dv = 0;
    if (c < W - 1) dh = xc - (double)x[f + 1];
    if (r < H - 1) dv = xc - (double)x[f + W];
    v[1] = dh * dh + dv * dv;
    // TV: c/mfista_TV_lib.c:30-46 -- isotropic inside, absolute on the last row/column
    if (c < W - 1 && r < H - 1)
      v[2] = sqrt(dh * dh + dv * dv);
    else if (c < W - 1)
      v[2] = fabs(dh);
    else if (r < H - 1)
      v[2] = fabs(dv);
  }
  block_sum<4>(v, red);
  if (threadIdx.x == 0)
    for (int q = 0; q < 4; q++) part[4 * blk + q] = v[q];
}

template <typename R>
__device__ inline void dev_momentum(int blk, R* __restrict__ x, const R* __restrict__ xnew, R* __restrict__ z, int n,
                                    const DevState* __restrict__ S) {
  if (S->gate_momentum) return;
  const int f = blk * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const R a = (R)S->mom_a, b = (R)S->mom_b;
  const R xf = x[f], xn = xnew[f];
  z[f] = xf * a + b * xn;
  if (S->accept_x) x[f] = xn;
}

// fixed-order sum of cnt doubles with stride `stride`, by the whole block
__device__ inline double block_reduce_range(const double* p, int cnt, int stride, double* red) {
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) v[0] += p[(size_t)i * stride];
  block_sum<1>(v, red);
  return v[0];  // valid in thread 0
}

// one workgroup: reduces a trial's partials, takes the backtracking and
// stopping decisions and sets the gates of the next step (red: >= 144 doubles)
__device__ inline void dev_control(DevState* S, double* cost, const double* __restrict__ part, PartialLayout L,
                                   double* red) {
  if (S->done) {
    if (threadIdx.x == 0) {
      S->gate_momentum = 1;
      S->gate_prologue = 1;
      S->gate_trial = 1;
    }
    return;
  }
  const bool prologue = (S->gate_prologue == 0);
  // all nine sums in one pass: thread t adds entries t, t + 256, ... of each
  // range in order, then one fixed-order block reduction -- the same sums,
  // bit for bit, as nine separate block_reduce_range calls
  double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  {
    const bool lip_on = S->lip_mode != 0;
    const int nmax = max(max(L.n_chiz, L.n_degrid), L.n_elem);
    for (int i = threadIdx.x; i < nmax; i += blockDim.x) {
      if (prologue && i < L.n_chiz) v[0] += part[L.chi_z + i];
      if (prologue && i < L.n_elem) v[1] += part[L.tsv_z + i];
      if (i < L.n_degrid) {
        v[2] += part[L.chi_x + i];
        if (lip_on) v[3] += part[L.lip + i];
      }
      if (i < L.n_elem) {
        const double* q = part + L.prox + (size_t)kProxW * i;
        v[4] += q[0];
        v[5] += q[1];
        v[6] += q[2];
        v[7] += q[3];
        v[8] += q[4];
      }
    }
    block_sum<9>(v, red);
  }
  if (threadIdx.x != 0) return;
  const double chi_z = v[0], tsv_z = v[1], chi_x = v[2], lip = v[3];
  const double s_dg = v[4], s_d2 = v[5], l1n = v[6], tsvn = v[7], tsvd = v[8];
  const bool tsv = S->lambda_tsv > 0.0;
  const int it = S->iter;
  if (prologue) {
    cost[it] = S->costtmp;
    S->Qcore = (chi_z + S->chi_add) / 2 + (tsv ? S->lambda_tsv * tsv_z : 0.0);
  }
  double F = (chi_x + S->chi_add) / 2 + (tsv ? S->lambda_tsv * tsvn : 0.0);
  const double Q = -s_dg + S->c * s_d2 / 2 + S->Qcore;
  S->last_F = F;
  S->last_Q = Q;
  S->trial += 1;
  S->total_trials += 1;
  if (!isfinite(F) || !isfinite(Q)) {  // non-finite objective: stop instead of backtracking forever
    S->status = 1;
    S->done = 1;
    S->ITER = it + 1;
    S->gate_prologue = 1;
    S->gate_trial = 1;
    S->gate_momentum = 1;
    return;
  }
  const double eta = (it == 0) ? S->eta_first : kEta;
  bool ok;
  if (S->lip_mode) {
    const double lhs = lip / 2 + (tsv ? S->lambda_tsv * tsvd : 0.0);
    const double rhs = S->c * s_d2 / 2;
    if (!isfinite(lhs)) {
      S->status = 1;
      S->done = 1;
      S->ITER = it + 1;
      S->gate_prologue = 1;
      S->gate_trial = 1;
      S->gate_momentum = 1;
      return;
    }
    ok = (lhs <= rhs);
  } else {
    ok = (F <= Q);
  }
  if (!ok && S->trial < S->maxiter) {
    S->c *= eta;
    S->gate_prologue = 1;
    S->gate_trial = 0;
    S->gate_momentum = 1;
    return;
  }
  if (!ok) S->c *= eta;
  if (S->trial > S->max_trials) S->max_trials = S->trial;
  S->c /= kEta;
  const double mu = S->mu;
  const double munew = (1 + sqrt(1 + 4 * mu * mu)) / 2;
  F += S->lambda_l1 * l1n;
  if (S->tv_mode) F += S->lambda_tv * tsvn;  // slot 3 holds TV(xnew) in TV mode
  bool stop = false;
  if (F < cost[it]) {
    S->costtmp = F;
    S->mom_a = (1 - mu) / munew;
    S->mom_b = 1 + ((mu - 1) / munew);
    S->accept_x = 1;
    S->l1_x = l1n;
  } else {
    S->mom_a = 1 - (mu / munew);
    S->mom_b = mu / munew;
    S->accept_x = 0;
    if (it > 1 && S->l1_x == 0) stop = true;
  }
  if (!stop && it >= kMinIter) {
    const double d = cost[it - kTD] - cost[it];
    stop = (S->rel_tol > 0) ? (d <= S->rel_tol * cost[it]) : (d < S->eps_abs);
  }
  // a target objective (timing at equal final objective): the iterate after
  // this iteration is xnew if accepted, else x
  if (!stop && S->target_F > 0 && (F < cost[it] ? F : cost[it]) <= S->target_F) stop = true;
  S->mu = munew;
  S->trial = 0;
  S->gate_momentum = 0;
  if (stop) {
    S->done = 1;
    S->ITER = it + 1;
  } else {
    S->iter = it + 1;
    if (S->iter >= S->maxiter) {
      S->done = 1;
      S->ITER = S->maxiter;
    }
  }
  S->gate_prologue = S->done ? 1 : 0;
  S->gate_trial = S->done ? 1 : 0;
}

// initial cost: cost0 = chi(x0)/2 + l1 |x0|_1 + l_tsv TSV(x0)
__device__ inline void dev_init_control(DevState* S, const double* __restrict__ part, PartialLayout L,
                                        const double* __restrict__ stats, int nsb, double* red) {
  const double chi = block_reduce_range(part + L.chi_x, L.n_degrid, 1, red);
  const double l1 = block_reduce_range(stats + 0, nsb, 4, red);
  const double tsv = block_reduce_range(stats + 1, nsb, 4, red);
  const double tv = block_reduce_range(stats + 2, nsb, 4, red);
  if (threadIdx.x != 0) return;
  double c0 = (chi + S->chi_add) / 2 + S->lambda_l1 * l1;
  if (S->tv_mode)
    c0 += S->lambda_tv * tv;
  else if (S->lambda_tsv > 0)
    c0 += S->lambda_tsv * tsv;
  S->costtmp = c0;
  S->l1_x = l1;
  S->mu = 1.0;
  S->iter = 0;
  S->trial = 0;
  S->done = 0;
  S->ITER = 0;
  S->max_trials = 0;
  S->total_trials = 0;
  S->gate_prologue = 0;
  S->gate_trial = 0;
  S->gate_momentum = 1;
}

}  // namespace mf
