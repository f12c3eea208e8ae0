/*
 * mfista.h -- C ABI of libmfista_amd.so, the MI355X (gfx950) MFISTA imaging solver.
 *
 * Drop-in for the reference library ikeda46/sparseimaging: the three
 * mfista_imaging_core_* entry points, their argument order and meaning, and
 * struct RESULT are those of c/mfista.h:45-66 and :151-178 (the C++ build,
 * c++/mfista.hpp:45-66 and :117-140, exports the same symbols plus
 * x2y_nufft / y2x_nufft).  All pointers of the reference entry points are
 * host pointers; the library uploads them, solves on the GPU and writes the
 * results back.  No torch or HIP types appear in these signatures.
 *
 * Errors follow the reference: invalid combinations print a message and
 * return with RESULT and xout unwritten.  The *_ex variants return a status
 * code instead (0 = success, < 0 = MFISTA_E_*).
 */
#ifndef MFISTA_AMD_H
#define MFISTA_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* algorithm constants, c/mfista.h:36-41 */
#define MAXITER 50000
#define MINITER 100
#define FGPITER 100
#define TD 50
#define ETA 1.1
#define EPS 1.0e-5

/* result record, field-for-field c/mfista.h:45-66 (128 bytes on LP64) */
struct RESULT {
  int M;
  int N;
  int NX;
  int NY;
  int N_active;
  int maxiter;
  int ITER;
  int nonneg;
  double lambda_l1;
  double lambda_tv;
  double lambda_tsv;
  double sq_error;
  double mean_sq_error;
  double l1cost;
  double tvcost;
  double tsvcost;
  double finalcost;
  double comp_time;
  double *residual;
  double Lip_const;
};

/* ---- reference entry points -------------------------------------------- */

/* replaces mfista_imaging_core_nufft, c/mfista.h:172-178 (c/mfista_nufft_lib.c:909);
 * NUFFT preset C-24 (24 taps), as the C library. */
void mfista_imaging_core_nufft(double *u_dx, double *v_dy, double *vis_r, double *vis_i,
                               double *vis_std, int M, int NX, int NY, int maxiter, double eps,
                               double lambda_l1, double lambda_tv, double lambda_tsv,
                               double cinit, double *xinit, double *xout, int nonneg_flag,
                               int box_flag, float *cl_box, struct RESULT *mfista_result);

/* replaces mfista_imaging_core_fft, c/mfista.h:161-168 (c/mfista_fft_lib.c:701);
 * fftw_plan_flag is accepted and ignored (there is no planner to tune). */
void mfista_imaging_core_fft(int *u_idx, int *v_idx, double *y_r, double *y_i,
                             double *noise_stdev, int M, int NX, int NY, int maxiter, double eps,
                             double lambda_l1, double lambda_tv, double lambda_tsv, double cinit,
                             double *xinit, double *xout, int nonneg_flag,
                             unsigned int fftw_plan_flag, int box_flag, float *cl_box,
                             struct RESULT *mfista_result);

/* replaces mfista_imaging_core_dft, c/mfista.h:151-157 (c/mfista_dft_lib.c:540);
 * A is M x N, column-major; looe_flag is accepted and ignored (as in the
 * reference, whose LOOE call is commented out).  RESULT.residual is malloc'ed
 * with M doubles and owned by the caller. */
void mfista_imaging_core_dft(double *y, double *A, int *M, int *N, int NX, int NY, int maxiter,
                             double eps, double lambda_l1, double lambda_tv, double lambda_tsv,
                             double cinit, double *xinit, double *xout, int nonneg_flag,
                             int looe_flag, int box_flag, float *cl_box,
                             struct RESULT *mfista_result);

/* replaces compute_LOOE_L1, c/looe_lib.c:195-216: the leave-one-out error
 * estimate of an L1 DFT-mode solution xvec of yvec ~ Amat x (Amat M x N,
 * column-major; yAx = yvec - Amat xvec).  Computed on the GPU in FP64.  Returns
 * the LOOE and sets *looe_m (= LOOE) and *looe_std, or returns -1 with both
 * set to 0 when the active-set Hessian is not positive definite (as the
 * reference does when dposv fails).  lambda1 and yvec are unused there too. */
double compute_LOOE_L1(int *M, int *N, double lambda1, double *yvec, double *Amat, double *xvec,
                       double *yAx, double *looe_m, double *looe_std);

/* replaces compute_LOOE_L1_TSV, c/looe_lib.c:276-309: as compute_LOOE_L1 with
 * lambda_tsv d2_TSV (4, 6 or 8 per pixel, c/looe_lib.c:218-235) added to the
 * Hessian's diagonal.  lambda_l1 and yvec are unused there too. */
double compute_LOOE_L1_TSV(int *M, int *N, int NX, int NY, double lambda_l1, double lambda_tsv,
                           double *yvec, double *Amat, double *xvec, double *yAx, double *looe_m,
                           double *looe_std);

/* replaces x2y_nufft, c++/mfista.hpp:136-137 (c++/mfista_nufft_lib.cpp:1076):
 * model visibilities y = A x with the C++ 12-tap preset, FP64. */
void x2y_nufft(double *u_dx, double *v_dy, int M, int Nx, int Ny, double *xin, double *y_r,
               double *y_i);

/* replaces y2x_nufft, c++/mfista.hpp:139-140 (c++/mfista_nufft_lib.cpp:1151):
 * x = Re(A^H y) with the C++ 12-tap preset, FP64. */
void y2x_nufft(double *u_dx, double *v_dy, int M, int Nx, int Ny, double *yin_r, double *yin_i,
               double *xout);

/* ---- extensions (no reference counterpart) ----------------------------- */

#define MFISTA_PRESET_C24 0   /* c/mfista_nufft_lib.c: MSP=12, round, tau=12/N^2 */
#define MFISTA_PRESET_CXX12 1 /* c++/mfista_nufft_lib.cpp: MSP=6, floor, tau=6/N^2, eta0=10 */

#define MFISTA_ENGINE_PERSISTENT 0
#define MFISTA_ENGINE_STREAM 1

/* The solve loop of a single NUFFT solve on the persistent engine (opts->loop).
 * GRIDDING: every iteration grids the residual and every trial degrids its
 * step, the reference's structure.  TOEPLITZ: the data term's normal operator
 * A^H W^2 A applied as a convolution with FP64 FFTs on the padded grid, its
 * kernel and A^H W^2 V computed once per solve with a 24-tap FP64 NUFFT
 * adjoint; no gridding inside the loop.  No counterpart in the reference. */
#define MFISTA_LOOP_AUTO 0
#define MFISTA_LOOP_GRIDDING 1
#define MFISTA_LOOP_TOEPLITZ 2

#define MFISTA_OK 0
#define MFISTA_E_BOTH_REG -1   /* lambda_tv and lambda_tsv both positive */
#define MFISTA_E_TV_PRESET -2  /* TV with the C++ preset (not in that library) */
#define MFISTA_E_FOV -3        /* a gridding tap leaves the 2NX x 2NY grid */
#define MFISTA_E_SIZE -4       /* unsupported image size */
#define MFISTA_E_ARG -5        /* invalid argument */
#define MFISTA_E_NONNEG -6     /* nonneg_flag not in {0,1}: ITER = 0, as the reference */
#define MFISTA_E_HIP -100      /* HIP runtime error */
#define MFISTA_E_COOP -101     /* the runtime refused the cooperative launch (another process
                                  holds the device's cooperative-launch resources); single
                                  solves then use the stream engine instead */

typedef struct mfista_opts {
  int preset;       /* MFISTA_PRESET_* (NUFFT only); default C24 */
  int fgp_iter;     /* FGP iterations of the TV prox; default FGPITER (100) */
  double rel_tol;   /* > 0: stop when cost[t-TD]-cost[t] <= rel_tol*cost[t], t >= MINITER;
                       0: the reference's absolute rule with eps (default) */
  int verbose;      /* nonzero: print progress */
  int max_trials;   /* out: largest number of backtracking trials in one iteration */
  double solve_ms;  /* out: GPU wall time of the solve (excluding the FP64 RESULT pass) */
  int total_trials; /* out: backtracking trials over the whole solve (forward NUFFTs of xnew) */
  int engine;       /* NUFFT solves: MFISTA_ENGINE_PERSISTENT (default, one cooperative
                       launch per solve) or MFISTA_ENGINE_STREAM (one launch per kernel) */
  double kernel_ms; /* out: persistent engine, duration of the solver launch (HIP events) */
  int trial_rounds; /* out: persistent engine, backtracking rounds evaluated; an L1+TSV round
                       evaluates the trials c and c*ETA together (MFISTA_PERSISTENT_PAIRED=0:
                       one trial per round) and takes them in order, as the reference does */
  int single_trials; /* in: nonzero = one backtracking trial per round (persistent engine) */
  int engine_used;   /* out: NUFFT solves, the engine that ran (MFISTA_ENGINE_*): a single
                        solve falls back to the stream engine when the runtime refuses the
                        cooperative launch */
  int deterministic; /* in: nonzero = reproducible NUFFT solves (persistent engine): every
                        gridding tile is summed by one workgroup in a fixed order, with no
                        floating-point atomics, so repeated solves are bit-identical */
  int autotune;      /* in: nonzero = NUFFT solves and batches (persistent engine) use the
                        launch configuration the autotuner chose for their shape, tuning
                        first (on the problem itself) when the table has none */
  int tune_iters;    /* in: iterations per autotuner candidate (default 30) */
  int loop;          /* in: NUFFT solves on the persistent engine: the Toeplitz loop applies
                        to L1 and L1+TSV square images of even size up to 512 x 512 (not with
                        deterministic, single_trials or the allocation-free entry point).
                        MFISTA_LOOP_AUTO (default) runs it where it applies (a shape the
                        autotuner timed runs its tuned loop), the gridding loop otherwise;
                        MFISTA_LOOP_TOEPLITZ runs it wherever it applies, also on a shape
                        tuned to the gridding loop; MFISTA_LOOP_GRIDDING always runs the
                        gridding loop.  Batches follow the same rule */
  int loop_used;     /* out: MFISTA_LOOP_GRIDDING or MFISTA_LOOP_TOEPLITZ, the loop that ran
                        (batches: the steps of the persistent or the wide engine); 0 when a
                        single solve ran on the stream engine */
} mfista_opts;

void mfista_default_opts(mfista_opts *opts);

/* same arguments as mfista_imaging_core_nufft plus options; returns a status */
int mfista_imaging_core_nufft_ex(const double *u_dx, const double *v_dy, const double *vis_r,
                                 const double *vis_i, const double *vis_std, int M, int NX,
                                 int NY, int maxiter, double eps, double lambda_l1,
                                 double lambda_tv, double lambda_tsv, double cinit,
                                 const double *xinit, double *xout, int nonneg_flag,
                                 int box_flag, const float *cl_box,
                                 struct RESULT *mfista_result, mfista_opts *opts);

/* Same as mfista_imaging_core_nufft_ex, but every array is a DEVICE pointer
 * (HBM-resident inputs, e.g. from hipMalloc or a torch tensor's data_ptr) and
 * the work is queued on `stream` (a hipStream_t; NULL = the HIP null stream),
 * so it is ordered after the caller's earlier work on that stream: inputs
 * written there need no synchronisation before the call.  The call returns
 * when the solve has finished and xout is written.  RESULT and opts stay host
 * structures. */
int mfista_nufft_solve_device(const double *d_u, const double *d_v, const double *d_vis_r,
                              const double *d_vis_i, const double *d_vis_std, int M, int NX,
                              int NY, int maxiter, double eps, double lambda_l1,
                              double lambda_tv, double lambda_tsv, double cinit,
                              const double *d_xinit, double *d_xout, int nonneg_flag,
                              int box_flag, const float *d_cl_box,
                              struct RESULT *mfista_result, mfista_opts *opts, void *stream);

/* One problem of a batch: the arguments of mfista_imaging_core_nufft (host
 * pointers).  No counterpart in the reference, which solves one problem per
 * call; a batch is a sequence of c/mfista_nufft_lib.c:909 calls. */
typedef struct {
  const double *u_dx, *v_dy, *vis_r, *vis_i, *vis_std;
  int M, NX, NY, maxiter;
  double eps, lambda_l1, lambda_tv, lambda_tsv, cinit;
  const double *xinit;
  double *xout;
  int nonneg_flag, box_flag;
  const float *cl_box;
  double target_objective; /* > 0: the solve also stops at the first iteration whose objective
                              is at or below this value (timing at equal final objective);
                              such batches run on the wide engine */
} mfista_nufft_problem;

/* Solves nprob independent NUFFT problems in one persistent launch (more
 * launches if there are more problems than workgroups); results[i] receives
 * problem i's RESULT exactly as mfista_imaging_core_nufft_ex would.  Problems
 * whose (u_dx, v_dy, vis_r, vis_i, vis_std, M, NX, NY) are the same pointers
 * and sizes share one NUFFT set-up (a lambda grid over one visibility set).
 * All problems use opts->preset and opts->fgp_iter, opts->rel_tol.  Returns 0
 * or the status of the first invalid problem (then nothing is solved); a
 * problem with nonneg_flag outside {0, 1} gets ITER = 0, as in the reference
 * (c/mfista_nufft_lib.c:634-637).  opts->solve_ms and opts->kernel_ms report
 * the whole batch.  Engines (opts->engine): MFISTA_ENGINE_PERSISTENT (default)
 * runs the batch in one cooperative launch, workgroups dealt to problems;
 * MFISTA_ENGINE_STREAM runs the wide engine when every problem is L1 or
 * L1+TSV of one image size (otherwise the persistent launch): the stream
 * engine's loop with every kernel launched once for all problems, one step
 * captured as a hipGraph; opts->engine_used reports which ran.  A batch with
 * a target_objective set runs on the wide engine (L1 / L1+TSV of one size,
 * otherwise MFISTA_E_ARG). */
int mfista_imaging_core_nufft_batch(int nprob, const mfista_nufft_problem *probs,
                                    struct RESULT *results, mfista_opts *opts);

/* Coupled movie: replaces matlab/MFISTA_movie_eht.m:1-239 (the reference's
 * only movie solver, a MATLAB function over FINUFFT), same arguments in the
 * same order and meaning.  The N_t = max(time_idx) frames are solved as one
 * problem: one step-size constant and one momentum sequence, objective
 *   sum_t |W_t (V_t - A_t x_t)|^2 / 2 + lambda_l1 sum_t |x_t|_1
 *     + lambda_tsv sum_t TSV(x_t) + lambda_t sum_t w_t |x_t - x_{t+1}|^2,
 * w_t = 1 / (t_{t+1} - t_t).
 *   time_idx[M]   frame of each visibility, 1-based (time_idx(:,1))
 *   time_val[M]   NULL, or observing time of each visibility (time_idx(:,2)):
 *                 t_t = N_t (mean time of frame t) / (max - min), else t_t = t
 *   c_box         NULL, or an NX*NY mask multiplied into xinit and the prox
 *                 output (a zero pixel stays zero; other values are taken as 1)
 *   xinit         NULL (zeros) or one NX*NY image used for every frame
 *   xout          N_t * NX * NY doubles, frame-major, x[t][i*NY + j]
 *   cost          NULL or maxiter doubles: the objective at the start of each
 *                 iteration (cost(1:t))
 *   terms         NULL or 4 doubles: sum_t |W_t (V_t - A_t x_t)|^2, sum |x|_1,
 *                 sum TSV, d_X (the values the reference collects in `error`)
 *   outvis_r/_i   NULL or M doubles: model visibilities A_t x_t (outvis)
 * eps is the reference's absolute rule cost[t-TD] - cost[t] < eps (unscaled);
 * opts->rel_tol > 0 selects the relative rule, opts->preset the gridding
 * kernel (default C-24, the closest to FINUFFT at 1e-12).  Returns a status. */
int mfista_movie_nufft(const double *vis_r, const double *vis_i, const double *u_dx,
                       const double *v_dy, const double *noise_std, const int *time_idx,
                       const double *time_val, int M, const float *c_box, int NX, int NY,
                       const double *xinit, double lambda_l1, double lambda_tsv, double lambda_t,
                       int nonneg_flag, double cinit, int maxiter, double eps, double *xout,
                       double *cost, double *terms, double *outvis_r, double *outvis_i,
                       int *iterations, mfista_opts *opts);

/* same arguments as mfista_imaging_core_fft (without fftw_plan_flag) plus options */
int mfista_imaging_core_fft_ex(const int *u_idx, const int *v_idx, const double *y_r,
                               const double *y_i, const double *noise_stdev, int M, int NX, int NY,
                               int maxiter, double eps, double lambda_l1, double lambda_tv,
                               double lambda_tsv, double cinit, const double *xinit, double *xout,
                               int nonneg_flag, int box_flag, const float *cl_box,
                               struct RESULT *mfista_result, mfista_opts *opts);

/* same arguments as mfista_imaging_core_dft (M, N by value, no looe_flag) plus options */
int mfista_imaging_core_dft_ex(const double *y, const double *A, int M, int N, int NX, int NY,
                               int maxiter, double eps, double lambda_l1, double lambda_tv,
                               double lambda_tsv, double cinit, const double *xinit, double *xout,
                               int nonneg_flag, int box_flag, const float *cl_box,
                               struct RESULT *mfista_result, mfista_opts *opts);

/* DFT-mode products for K right-hand sides that share A, on the FP64 matrix
 * cores: R_k = y - A x_k (y NULL: zero) and, if G is not NULL, G_k = A^T R_k.
 * The reference computes one dgemv per vector (c/mfista_dft_lib.c:30-48,
 * calc_yAz / dL_dx); with many vectors the products are GEMMs.  A is M x N,
 * column-major; X holds K vectors of N (X[k N + n]), R K vectors of M, G K
 * vectors of N; host pointers; 1 <= K <= 64.  Returns a status. */
int mfista_dft_products(const double *A, int M, int N, int K, const double *X, const double *y,
                        double *R, double *G);

/* nprob DFT-mode problems that share y and A and differ in lambda_l1,
 * lambda_tsv and cinit (a regularisation grid): each is solved exactly as
 * mfista_imaging_core_dft_ex would solve it (the loop of
 * c/mfista_dft_lib.c:282-436, L1 or L1+TSV, eps unscaled, opts->rel_tol
 * honoured), but up to 64 problems advance together and their products with A
 * are GEMMs on the FP64 matrix cores (mfista_dft_products).  xinit: NULL
 * (zeros) or nprob images of N = NX NY values (xinit[k N + n]); xout: nprob
 * images; results[k] as mfista_imaging_core_dft fills it (residual malloc'ed
 * with M doubles, owned by the caller).  One cl_box (box_flag = 1) applies to
 * every problem.  TV is not offered here (lambda_tv is not an argument).
 * No counterpart in the reference, which solves one problem per call. */
int mfista_imaging_core_dft_batch(const double *y, const double *A, int M, int N, int NX, int NY,
                                  int nprob, const double *lambda_l1, const double *lambda_tsv,
                                  const double *cinit, int maxiter, double eps, const double *xinit,
                                  double *xout, int nonneg_flag, int box_flag, const float *cl_box,
                                  struct RESULT *results, mfista_opts *opts);

/* the LOOE of compute_LOOE_L1_TSV (lambda_tsv = 0: compute_LOOE_L1) with a
 * status return: *hessian_positive = 1 and *looe_m, *looe_std set, or 0 (both
 * 0) when the Hessian is not positive definite; *n_active (may be NULL)
 * receives the size of the active set.  M >= 2. */
int mfista_looe_dft(int M, int N, int NX, int NY, double lambda_tsv, const double *Amat,
                    const double *xvec, const double *yAx, double *looe_m, double *looe_std,
                    int *hessian_positive, int *n_active);

/* standalone NUFFT operators with a chosen preset and precision (0 = FP32,
 * 1 = FP64); host pointers; return a status */
int mfista_nufft_forward(int preset, int fp64, const double *u, const double *v, int M, int NX,
                         int NY, const double *x, double *y_r, double *y_i);
int mfista_nufft_adjoint(int preset, int fp64, const double *u, const double *v, int M, int NX,
                         int NY, const double *y_r, const double *y_i, double *x);

/* Tracing hook: re-runs each kernel of the last NUFFT solve `reps` times on
 * the library stream with one HIP event pair per launch and writes the mean
 * duration in ms of stage s to ms[s]: 0 forward row FFT, 1 forward column
 * FFT, 2 degridding, 3 gridding, 4 adjoint column FFT, 5 adjoint row c2r,
 * 6 TSV gradient, 7 prox.  Returns a status. */
int mfista_profile_stages(int reps, double *ms, int nstage);

/* number of trial steps the last NUFFT solve enqueued */
long mfista_last_steps(void);

const char *mfista_amd_version(void);

/* Allocation-free variant of mfista_nufft_solve_device: the same arguments,
 * the same solve and results, but every workspace comes from a scratch array
 * of MFISTA_NOALLOC_SCRATCH_BYTES that the library's code object holds on each
 * device (from the first call on that device until the process ends), so the
 * call makes no hipMalloc/hipFree and creates no stream, event or pinned
 * buffer.  No counterpart in the reference: for callers that cannot allocate
 * device memory (tight serving loops, harnesses that forbid allocation).
 * A problem whose workspaces do not fit returns MFISTA_E_SIZE with a message
 * on stderr (a 256x256 solve with M = 20000 uses well under half of it).  Only the persistent
 * engine runs here (opts->engine must be MFISTA_ENGINE_PERSISTENT; a refused
 * cooperative launch returns MFISTA_E_COOP), and opts->kernel_ms is not
 * measured (0).  All calls on one device share its scratch array: the library
 * serialises them and each returns when its solve has finished; calls that
 * overlapped on one device would race. */
#define MFISTA_NOALLOC_SCRATCH_BYTES (64u << 20)
int mfista_nufft_solve_device_noalloc(const double *d_u, const double *d_v, const double *d_vis_r,
                                      const double *d_vis_i, const double *d_vis_std, int M, int NX,
                                      int NY, int maxiter, double eps, double lambda_l1,
                                      double lambda_tv, double lambda_tsv, double cinit,
                                      const double *d_xinit, double *d_xout, int nonneg_flag,
                                      int box_flag, const float *d_cl_box,
                                      struct RESULT *mfista_result, mfista_opts *opts, void *stream);

/* Diagnostic: the number of device allocations (hipMalloc) the library has
 * made since it was loaded. */
long mfista_device_allocations(void);

/* ---- launch-configuration autotuner --------------------------------------
 * No counterpart in the reference, whose schedule is fixed (its nearest
 * mechanism is configure_tile, c++/mfista_nufft_lib.cpp:643-678, which
 * balances gridding taps over OpenMP threads by rule).  The persistent
 * solver's gridding tile (16x16, 32x16 or 32x32 cells), workgroups per CU,
 * paired backtracking rounds, the FGP block shape (L1+TV) and, for batches,
 * the engine are timed on the problem itself: every candidate
 * solves it for `iters` iterations from its xinit (no stopping rule), twice,
 * and the fastest per iteration is kept in a process-wide table under the
 * shape key (NX, NY, preset, regulariser class, floor(log2 M), floor(log2
 * problems)).  Solves and batches with opts->autotune set use the table's
 * entry for their shape and tune on a miss.  MFISTA_TUNE_CACHE=<file> loads
 * a table saved earlier when the library starts and rewrites the file after
 * every new entry. */
typedef struct mfista_tune {
  int NX, NY, preset; /* key: image size and gridding preset */
  int reg;            /* key: 0 L1, 1 L1+TSV, 2 L1+TV */
  int m_log2;         /* key: floor(log2 M) (the largest M of a batch) */
  int nprob_log2;     /* key: floor(log2 problems): 0 for a single solve */
  int tile_x, tile_y; /* gridding tile, cells along kx and ky */
  int blocks_per_cu;  /* workgroups per CU of the cooperative launch */
  int paired;         /* paired backtracking rounds (L1 and L1+TSV) */
  int engine;         /* batches: MFISTA_ENGINE_PERSISTENT (one cooperative launch) or
                         MFISTA_ENGINE_STREAM (the wide engine: every stage one launch for
                         all problems, a step replayed as a graph) */
  int fgp_tw;         /* L1+TV: FGP block shape, 0 = pixels one per thread along the flat
                         image, 64 / 32 / 16 = tiles of fgp_tw x (256 / fgp_tw) pixels */
  double us_per_iter; /* measured microseconds per iteration */
  int loop;           /* single solves: MFISTA_LOOP_GRIDDING or MFISTA_LOOP_TOEPLITZ (the loop
                         the entry was timed on and that tuned solves run), 0 = by rule */
} mfista_tune;

/* Tunes the shape of this problem (arguments as mfista_imaging_core_nufft_ex,
 * without maxiter, eps, xout and the box; opts: preset, fgp_iter,
 * deterministic; NULL = defaults) and stores the choice.  iters <= 0: 30.
 * *best receives the choice; all (may be NULL) up to max_all candidates with
 * their times, *n_all (may be NULL) their number.  Returns a status. */
int mfista_autotune_nufft(const double *u_dx, const double *v_dy, const double *vis_r,
                          const double *vis_i, const double *vis_std, int M, int NX, int NY,
                          double lambda_l1, double lambda_tv, double lambda_tsv, double cinit,
                          const double *xinit, int nonneg_flag, const mfista_opts *opts, int iters,
                          mfista_tune *best, mfista_tune *all, int max_all, int *n_all);
/* The same for a batch (mfista_imaging_core_nufft_batch's problems; every
 * problem runs `iters` iterations; xout is not written). */
int mfista_autotune_nufft_batch(int nprob, const mfista_nufft_problem *probs,
                                const mfista_opts *opts, int iters, mfista_tune *best,
                                mfista_tune *all, int max_all, int *n_all);
/* the table: entries copied to out (up to cap); returns their number */
int mfista_tune_table(mfista_tune *out, int cap);
/* writes the table to a text file / merges a file into it (returns the number
 * of entries read); status < 0 on error */
int mfista_tune_save(const char *path);
int mfista_tune_load(const char *path);
void mfista_tune_clear(void);
/* candidate solves the autotuner has timed in this process */
long mfista_tune_runs(void);

/* Releases the library's per-device state: waits for its work, frees its
 * device and pinned host memory, destroys its streams and events.  Optional;
 * call it before the HIP runtime is torn down (e.g. at the end of main, or from
 * an atexit handler registered after the first call), not from a static
 * destructor.  A later call of any entry point starts afresh.  No counterpart
 * in the reference, which keeps no state between calls. */
void mfista_amd_finalize(void);

#ifdef __cplusplus
}
#endif
#endif
