// driver.h -- the MFISTA driver shared by the NUFFT, FFT and DFT modes.
#pragma once

#include "solver.h"

namespace mf {

// A data-term operator.  fwd(x) writes per-block partial sums whose total is
// 2 F(x) (F = chi^2 / 2 in NUFFT/DFT mode, the half-plane norm / 4 in FFT
// mode); with keep = true it also keeps what adj() needs.  adj(g) writes
// g = -dF/dx at the last kept point.  Every launch is gated by `skip`.
template <typename R>
struct Operator {
  virtual ~Operator() = default;
  virtual int npart() const = 0;
  virtual void fwd(const R* x, bool keep, double* part, const int* skip) = 0;
  virtual void adj(R* g, const int* skip) = 0;
  // optional: forward of a step d from the kept point; partials of 2 F(z + d)
  // and of |W A d|^2 (enables the cancellation-free backtracking test)
  virtual bool has_delta() const { return false; }
  virtual void fwd_delta(const R*, double*, double*, const int*) {}
};

struct RunParams {
  int NX = 0, NY = 0, maxiter = 0, nonneg = 0, box_flag = 0, tv_mode = 0, fgp_iter = 100;
  double eps_abs = 0, rel_tol = 0, lambda_l1 = 0, lambda_tv = 0, lambda_tsv = 0, cinit = 0;
  double eta_first = 1.1;
  int single_trials = 0;          // persistent engine: one backtracking trial per round
  int deterministic = 0;          // persistent engine: gridding without atomics (tile-exact ranges)
  double target_F = 0;            // stream and wide engines: > 0 also stops at this objective
  int fgp_tw = 0;                 // persistent engine: FGP block shape (0 = along the flat image,
                                  // else tiles of fgp_tw x 256 / fgp_tw pixels)
  const double* xinit = nullptr;  // device, FP64
  double* xout = nullptr;         // device, FP64
  const float* box = nullptr;     // device
};

struct RunStats {
  int ITER = 0, max_trials = 0, total_trials = 0, status = 0;
  double c = 0;
  double kernel_ms = 0;  // persistent engine: duration of the solver launch
  long steps = 0;
  int rounds = 0;        // persistent engine: trial rounds (one or two trials each)
};

template <typename R>
int mfista_run(Context& C, Operator<R>& op, const RunParams& P, RunStats* out);

// the whole NUFFT solve in one cooperative launch (persistent.hip); returns
// kCoopRefused, with nothing run, if the runtime refuses the cooperative launch
constexpr int kCoopRefused = MFISTA_E_COOP;
int solve_nufft_persistent(Context& C, const RunParams& P, RunStats* out);
// the Toeplitz loop (toeplitz_body.h): its set-up for the problem of C.nu
// (toeplitz.hip; MFISTA_E_SIZE or MFISTA_E_FOV when it does not apply) and
// the solve in one cooperative launch (persistent.hip)
int toeplitz_setup(Context& C, const double* d_u, const double* d_v, const double* d_vr, const double* d_vi,
                   const double* d_sd, int M, int NX, int NY, int preset, double wv2, ToepOut& out);
int solve_nufft_toeplitz(Context& C, const RunParams& P, RunStats* out);
// independent Toeplitz-loop solves in one cooperative launch: problem i uses
// sets[i] (its NUFFT set-up, for the geometry) and ops[i] (its set's operator
// data from toeplitz_setup), C.bwork[i] its workspace (persistent.hip)
int solve_nufft_toeplitz_batch(Context& C, const std::vector<NufftState*>& sets,
                               const std::vector<const ToepOut*>& ops, const std::vector<RunParams>& P,
                               std::vector<RunStats>& out, float* kernel_ms);
int solve_nufft_batch(Context& C, const std::vector<NufftState*>& sets, const std::vector<RunParams>& P,
                      std::vector<RunStats>& out, float* kernel_ms);
// independent solves of one image size and preset with every kernel launched
// once for all problems, one step captured as a graph (wide.hip); L1 and
// L1+TSV only.  ops (square images): the Toeplitz steps, with problem i's
// visibility set's operator data ops[i] (toeplitz_setup)
int solve_nufft_wide(Context& C, const std::vector<NufftState*>& sets, const std::vector<RunParams>& P,
                     std::vector<RunStats>& out, float* kernel_ms,
                     const std::vector<const ToepOut*>* ops = nullptr);
// the frames of a coupled movie in one launch with one control step (persistent.hip)
int solve_nufft_movie(Context& C, const std::vector<NufftState*>& sets, const std::vector<RunParams>& P,
                      double lambda_t, const std::vector<double>& w, std::vector<RunStats>& out,
                      float* kernel_ms);

// FGP TV prox kernels (reference c/mfista_TV_lib.c:189-384)
template <typename R>
void launch_tv_pre(const R* z, const R* g, R* b, int n, const DevState* S, hipStream_t st,
                   const int* skip);
template <typename R>
void launch_fgp_iter(const R* b, const float* box, const R* rin, const R* sin, R* rout, R* sout,
                     R* p, R* q, int W, int H, const DevState* S, double c1, double c2, int first,
                     hipStream_t st, const int* skip);
template <typename R>
void launch_fgp_final(const R* b, const float* box, const R* p, const R* q, const R* z,
                      const R* g, R* xnew, R* d, int W, int H, const DevState* S, double* part,
                      int have_dual, hipStream_t st, const int* skip);

}  // namespace mf
