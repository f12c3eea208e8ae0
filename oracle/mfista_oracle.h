/*
 * mfista_oracle.h -- FP64 CPU restatement of the reference MFISTA imaging library.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the GPU
 * solver in sparseimaging_amd.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never links it.
 *
 * Every function restates one reference function; the comment above each
 * definition in mfista_oracle.c cites the reference file:line it follows.
 * Reference: ikeda46/sparseimaging, files c/ and c++/ (GPL-3); nothing is
 * copied, the behaviour is restated.
 *
 * Two NUFFT presets:
 *   OR_PRESET_C24   c/mfista_nufft_lib.c    MSP=12 (24 taps), tau=12/N^2, round()
 *   OR_PRESET_CXX12 c++/mfista_nufft_lib.cpp MSP=6 (12 taps), tau=6/N^2, floor(),
 *                   points with v<0 flipped, first-iteration backtracking factor 10
 */
#ifndef MFISTA_ORACLE_H
#define MFISTA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* algorithm constants: c/mfista.h:36-41 */
#define OR_MAXITER 50000
#define OR_MINITER 100
#define OR_FGPITER 100
#define OR_TD 50
#define OR_ETA 1.1

#define OR_PRESET_C24 0
#define OR_PRESET_CXX12 1

/* Same field order and types as struct RESULT, c/mfista.h:45-66 */
typedef struct or_result {
  int M, N, NX, NY, N_active, maxiter, ITER, nonneg;
  double lambda_l1, lambda_tv, lambda_tsv, sq_error, mean_sq_error, l1cost,
      tvcost, tsvcost, finalcost, comp_time;
  double *residual;
  double Lip_const;
} or_result;

typedef struct or_opts {
  int preset;      /* OR_PRESET_* (NUFFT only) */
  int fgp_iter;    /* FGP inner iterations; reference constant is FGPITER=100 */
  double rel_tol;  /* >0: stop when cost[t-TD]-cost[t] <= rel_tol*cost[t] (t>=MINITER);
                      0: the reference's absolute rule with the scaled eps */
  int verbose;     /* print progress like the reference (every 10/100 iterations) */
  double *cost_trace; /* optional: receives cost[0..ITER-1] (length >= maxiter) */
  int max_trials_out; /* filled: largest backtracking trial count seen */
} or_opts;

void or_default_opts(or_opts *o);

/* ---- FFT (unnormalised, FFTW sign convention: forward e^{-2 pi i}) ---- */
void or_fft_c2c(int n, double *data /* n complex, interleaved */, int sign);
void or_fft2_r2c(int n0, int n1, const double *in, double *out /* n0*(n1/2+1) cplx */);
void or_fft2_c2r(int n0, int n1, const double *in /* n0*(n1/2+1) cplx */, double *out);

/* ---- elementwise tools: c/mfista_tools.c ---- */
double or_calc_Q_part(int N, const double *x1, const double *x2, double c,
                      const double *AyAz, double *buf);
void or_soft_threshold_box(const double *v, int n, double eta, double *out,
                           int box_flag, const float *cl_box);
void or_soft_threshold_nonneg_box(const double *v, int n, double eta, double *out,
                                  int box_flag, const float *cl_box);
double or_TSV(int NX, int NY, const double *x);
void or_d_TSV(int NX, int NY, const double *x, double *d);

/* ---- TV / FGP: c/mfista_TV_lib.c ---- */
double or_TV(int NX, int NY, const double *x);
void or_Lx(int NX, int NY, const double *x, double *p, double *q);
void or_Lpq(int NX, int NY, const double *p, const double *q, double *x);
void or_P_pqiso(int NX, int NY, const double *p, const double *q, double *r, double *s);
void or_FGP_L1_box(int NX, int NY, const double *b, double lambda_l1, double lambda_tv,
                   int ITER, double *x, int box_flag, const float *cl_box);
void or_FGP_nonneg_box(int NX, int NY, const double *b, double lambda_tv, int ITER,
                       double *x, int box_flag, const float *cl_box);

/* ---- NUFFT operators (A.x = NUFFT2d2, Re(A^H r) = NUFFT2d1) ---- */
/* returns 0, or -1 when a tap would leave the 2N x 2N grid (FOV violation) */
int or_nufft_forward(int preset, const double *u, const double *v, int M, int Nx, int Ny,
                     const double *x, double *yr, double *yi);
int or_nufft_adjoint(int preset, const double *u, const double *v, int M, int Nx, int Ny,
                     const double *yr, const double *yi, double *x);

/* ---- FFT-mode objective pieces: c/mfista_fft_lib.c ---- */
/* F = 1/4 sum_full |mask (y - FFT(x)/sqrt(NN))|^2 and g = -dF/dx */
double or_fft_F_grad(int M, int NX, int NY, const int *u_idx, const int *v_idx,
                     const double *y_r, const double *y_i, const double *noise,
                     const double *x, double *g /* may be NULL */);

/* ---- top-level solvers, same arguments as c/mfista.h:151-178 plus options ---- */
int or_imaging_core_nufft(const double *u_dx, const double *v_dy, const double *vis_r,
                          const double *vis_i, const double *vis_std, int M, int NX, int NY,
                          int maxiter, double eps, double lambda_l1, double lambda_tv,
                          double lambda_tsv, double cinit, const double *xinit, double *xout,
                          int nonneg_flag, int box_flag, const float *cl_box,
                          or_result *res, or_opts *opts);
int or_imaging_core_fft(const int *u_idx, const int *v_idx, const double *y_r,
                        const double *y_i, const double *noise_stdev, int M, int NX, int NY,
                        int maxiter, double eps, double lambda_l1, double lambda_tv,
                        double lambda_tsv, double cinit, const double *xinit, double *xout,
                        int nonneg_flag, int box_flag, const float *cl_box, or_result *res,
                        or_opts *opts);
int or_imaging_core_dft(const double *y, const double *A, int M, int N, int NX, int NY,
                        int maxiter, double eps, double lambda_l1, double lambda_tv,
                        double lambda_tsv, double cinit, const double *xinit, double *xout,
                        int nonneg_flag, int box_flag, const float *cl_box, or_result *res,
                        or_opts *opts);

/* objective of a given image, as calc_result_* computes it (FP64) */
int or_result_nufft(int preset, const double *u, const double *v, const double *vis_r,
                    const double *vis_i, const double *vis_std, int M, int NX, int NY,
                    double lambda_l1, double lambda_tv, double lambda_tsv, const double *x,
                    or_result *res);

void or_free(void *p);

#ifdef __cplusplus
}
#endif
#endif
