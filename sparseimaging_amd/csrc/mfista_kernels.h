// mfista_kernels.h -- MFISTA outer-loop kernels and the device-resident state.
//
// The loop of the reference (c/mfista_nufft_lib.c:639-752, identical in the
// FFT and DFT libraries) runs as a device state machine: every enqueued "step"
// is one backtracking trial, preceded by the iteration prologue (A z, A^H r,
// TSV gradient) when the previous trial was accepted.  A one-workgroup control
// kernel reduces the FP64 partial sums, applies the F <= Q test, updates c, mu
// and the cost history in FP64 and writes the gates that make the next step's
// kernels run or return at once.  The host only polls `done` every few steps.
//
// The backtracking test F(xnew) <= Q(xnew, z) of the reference is, because the
// adjoint used for the gradient is the exact adjoint of the forward operator,
// algebraically the same as  1/2 |W A d|^2 + lambda_tsv TSV(d) <= c/2 |d|^2
// with d = xnew - z (TSV is quadratic).  With FP32 operators the second form
// is used (lip_mode = 1): it is a comparison of two sums of squares, free of
// the cancellation between chi^2(xnew) and chi^2(z) that makes the first form
// flip on rounding noise near convergence.  FP64 operators use the reference
// form unchanged (lip_mode = 0).
#pragma once

#include "common.h"

namespace mf {

constexpr int kMinIter = 100;  // MINITER, c/mfista.h:37
constexpr int kTD = 50;        // TD, c/mfista.h:39
constexpr double kEta = 1.1;   // ETA, c/mfista.h:40

struct DevState {
  double c, mu, costtmp, Qcore, lambda_l1, lambda_tsv, lambda_tv, eps_abs, rel_tol, eta_first;
  double mom_a, mom_b, l1_x, last_F, last_Q;
  int iter, trial, done, ITER, maxiter, nonneg, tv_mode, max_trials, accept_x, status;
  int total_trials, lip_mode;
  int gate_prologue, gate_trial, gate_momentum;  // nonzero: the kernel returns at once
  int rounds;                                    // persistent engine: trial rounds evaluated
  double target_F;  // > 0: also stop once the objective is at or below it (stream and wide engines)
  double chi_add;   // added to every chi^2 sum: |W V|^2 in the Toeplitz loop (whose partials
                    // are <x, T x> - 2 <x, b>), 0 otherwise
};

// offsets (in doubles) of the partial-sum regions written by the kernels
constexpr int kProxW = 6;  // per-block partials of the trial kernel
struct PartialLayout {
  int chi_z, chi_x, lip, tsv_z, prox;  // prox: kProxW values per block
  int n_degrid, n_elem, total;
  int n_chiz;  // partials of the chi^2 at z (n_degrid unless a separate kernel writes them)
};

template <typename R>
void launch_tsv_grad(const R* z, R* g, int W, int H, double lambda_tsv, double* part,
                     hipStream_t st, const int* skip);
template <typename R>
void launch_prox(const R* z, const R* g, R* xnew, R* d, const float* box, int W, int H,
                 const DevState* S, double* part, hipStream_t st, const int* skip);
template <typename R>
void launch_stats(const R* x, int W, int H, double* part, hipStream_t st);  // 4 per block
template <typename R>
void launch_momentum(R* x, const R* xnew, R* z, int n, const DevState* S, hipStream_t st);
void launch_control(DevState* S, double* cost, const double* part, PartialLayout L,
                    hipStream_t st);
void launch_init_control(DevState* S, const double* part, PartialLayout L, const double* stats,
                         int nstat_blocks, hipStream_t st);
template <typename Rs, typename Rd>
void launch_cast(const Rs* a, Rd* b, int n, hipStream_t st, const int* skip = nullptr);
int elem_blocks(int n);

}  // namespace mf
