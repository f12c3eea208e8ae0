// solver.h -- host-side context and workspaces shared by the solver files.
#pragma once

#include <vector>

#include <mutex>

#include "../../include/mfista.h"
#include "fft_lds.h"
#include "mfista_kernels.h"
#include "nufft.h"

namespace mf {

// grow-only device allocation; inside an ArenaScope it is carved from the
// scope's preallocated device buffer instead (no hipMalloc)
struct DevMem {
  void* p = nullptr;
  size_t n = 0;
  bool in_arena = false;  // carved from an arena: release() does not free it
  int ensure(size_t bytes);
  void release();
};

// A bump allocator over preallocated device memory.  While an ArenaScope is
// alive (library mutex held), every DevMem::ensure carves from it; a request
// that does not fit fails (and sets `exhausted`) instead of allocating.
struct Arena {
  char* base = nullptr;
  size_t cap = 0, off = 0, peak = 0;
  bool exhausted = false;
};
struct ArenaScope {
  explicit ArenaScope(Arena* a);
  ~ArenaScope();
};

// device allocations made by DevMem::ensure since the library was loaded
long device_allocations();

struct NufftState {
  bool ready = false, have64 = false;
  int preset = -1, NX = 0, NY = 0, M = 0, MSP = 0;
  GridGeom geom;
  FFTPlan Px, Py;
  DevMem twxf, twyf, twxd, twyd, e4xf, e4yf, e4xd, e4yd;
  DevMem mxy, af, e1f, visf, wf, ad, e1d, visd, wd, perm, yneg;
  DevMem keys_a, keys_b, idx_a, vals_a, ent, tstart, tend, sort_tmp, err, ones, zero_vis, chunks;
  int nchunks = 0;
  std::vector<int4> h_chunks;  // host copy of the chunk list
  DevMem T, Gf, Hg, U, rw, T64, Gf64, Hg64, U64, rw64;
  VisParams<float> pf() {
    return {(int2*)mxy.p, (float2*)af.p, (float*)e1f.p, (float2*)visf.p, (float*)wf.p};
  }
  VisParams<double> pd() {
    return {(int2*)mxy.p, (double2*)ad.p, (double*)e1d.p, (double2*)visd.p, (double*)wd.p};
  }
};

// Device workspace of one persistent-engine solve.
struct PersistWork {
  DevMem px, pxn, pz;      // FP64 iterates x, xnew, z
  DevMem g;                // FP32 gradient
  DevMem vmod;             // three FP64 weighted model vectors
  DevMem T64, Gf64;        // FP64 forward transform of the step
  DevMem Hg, U;            // FP32 adjoint grid and column-pass output
  DevMem part, state, cost, bar, fgp, fgpc, granges;
  DevMem pchunks;          // the gridding chunk list re-cut at the workgroups' range boundaries
  DevMem x0, xo, box;      // batched host-pointer solves: staged xinit, xout, cl_box
  DevMem pargs;            // the solver kernel's arguments (device copy)
  DevMem tvec, u64, xrot;  // Toeplitz loop: T applied to x / xnew / z, column-pass outputs, image buffers
};

// The Toeplitz loop's operator data of one visibility set (toeplitz.hip).
struct ToepOut {
  DevMem kh, bimg;  // the kernel's spectrum Kh, b = Re A^H W^2 V
  double wv2 = 0;   // |W V|^2
  FFTPlan Pxh;                 // the NX-point plan of the pruned row and column transforms
  DevMem twxh32, twxh64;       // and its twiddles
};
// The Toeplitz loop's set-up scratch and the single solve's operator data.
struct ToepWork {
  NufftState nu;                                // the set-up's 2L x 2L, 24-tap FP64 transform
  DevMem rb, rp, imgp, imgb, circ, tbuf, fbuf;  // set-up scratch
  ToepOut out;                                  // a single solve's operator data
};

// One visibility set of a batch, with its NUFFT set-up.
struct BatchSet {
  DevMem u, v, vr, vi, sd;
  NufftState nu;
  double vis_power = 0;  // sum |V|^2 (eps scaling)
  ToepOut toep;          // the Toeplitz loop's operator data for this set
};

// Device workspace of one problem of the wide engine (wide.hip).
struct WideWork {
  DevMem tvec;               // Toeplitz steps: T x, T xnew, T z
  DevMem x, z, xnew, gr, d;  // FP64 iterates, gradient, step
  DevMem T64, Gf64, mod;     // FP64 forward transform; models W A x, W A xnew, W A z
  DevMem rz32, Hg, U;        // FP32 adjoint
  DevMem part, stat, state, cost;
};

// Device workspace of the DFT-mode multi-problem solver (dft_batch.hip).
struct DftBatchBufs {
  DevMem A, y, X, Z, XN, D, G, Rx, Rz, Rn, part, psum, psq_n, psq_z, S, cost, box;
};

struct Context {
  int device = -1;
  bool no_alloc = false;  // the allocation-free entry point: workspaces from its arena, no events
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int* hflags = nullptr;  // pinned
  long last_steps = 0;
  NufftState nu;
  // solver workspaces
  DevMem x, z, xnew, g, dstep, x64, part, stat, stat64, chi64, state, cost;
  DevMem fgp;   // TV prox duals
  DevMem fgpc;  // FGP momentum coefficients
  PersistWork pw;  // the persistent solver's workspace (single solves)
  ToepWork tw;     // the Toeplitz loop's set-up and workspace (single solves)
  // batched solves: visibility sets (each with its NUFFT set-up) and one
  // workspace per problem
  std::vector<BatchSet> bset;
  std::vector<PersistWork> bwork;
  DevMem batch_args, batch_blk;  // PArgs array and workgroup table of a batch launch
  DevMem movie_part;             // partial-sum table shared by the frames of a coupled movie
  DevMem trace, trace_tag;       // persistent-engine barrier trace (MFISTA_PERSISTENT_TRACE)
  DevMem g32;          // stream engine: FP32 gradient before its FP64 copy
  DftBatchBufs dbb;    // DFT mode, many problems sharing one matrix
  std::vector<WideWork> wwork;   // wide engine: one workspace per problem
  DevMem wdesc, wdone;           // wide engine: problem descriptors, finished-problem counter
  hipEvent_t ev_t[2] = {nullptr, nullptr};  // timing events (wide engine)

  int persist_blocks_per_cu = 2;  // MFISTA_PERSISTENT_BLOCKS_PER_CU overrides
  int persist_blocks = 0;  // workgroups of the last persistent launch
  int persist_hier_barrier = 1;  // MFISTA_PERSISTENT_FLAT_BARRIER=1 selects the flat barrier
  int persist_pair = 1;          // MFISTA_PERSISTENT_PAIRED=0: one backtracking trial per round
  int grid_tile_x = 16, grid_tile_y = 16;  // gridding tile of the next NUFFT set-up (autotuner)
  int def_blocks_per_cu = 2, def_pair = 1;  // the knobs' defaults (environment overrides included)
  int tune_engine = -1;  // batches: the autotuner's engine (-1: the caller's opts->engine)
  int fgp_tw = 0;        // FGP block shape of the persistent engine (autotuner; 0 = along the flat image)
  int tune_loop = 0;     // single solves: the autotuner's loop (MFISTA_LOOP_*; 0 = by rule)
  // uploads of the host-pointer entry points
  DevMem h_u, h_v, h_vr, h_vi, h_sd, h_x0, h_x1, h_box, h_model, h_A, h_y, h_idx;
};

struct SolveArgs {
  const double *u = nullptr, *v = nullptr, *vr = nullptr, *vi = nullptr, *sd = nullptr;
  int M = 0, NX = 0, NY = 0, maxiter = 0, nonneg = 0, box_flag = 0;
  double eps = 0, eps_abs = 0, lambda_l1 = 0, lambda_tv = 0, lambda_tsv = 0, cinit = 0;
  const double* xinit = nullptr;
  double* xout = nullptr;
  const float* box = nullptr;
};

Context* get_context();
std::mutex& global_mutex();
int upload(Context& C, DevMem& m, const void* h, size_t bytes);
int nufft_setup(Context& C, int preset, const double* d_u, const double* d_v, const double* d_vr,
                const double* d_vi, const double* d_sd, int M, int NX, int NY);
int nufft_ensure64(Context& C);
template <typename R>
void nufft_fwd(Context& C, const R* x, cplx<R>* rz, double* chi_part, cplx<R>* model,
               const int* skip, float2* rz32 = nullptr);
template <typename R>
void nufft_fwd_delta(Context& C, const R* d, double* chi_part, double* lip_part, const int* skip);
template <typename R>
void nufft_adj(Context& C, const cplx<R>* rz, R* g, const int* skip);
int image_stats64(Context& C, const double* d_x64, int NX, int NY, double out[4]);
int nufft_chi2_64(Context& C, const double* d_x64, double* chi2);
int solve_nufft(Context& C, const SolveArgs& A, RESULT* res, mfista_opts* opts);

}  // namespace mf
