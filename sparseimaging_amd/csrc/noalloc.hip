// noalloc.hip -- the allocation-free variant of the device-pointer NUFFT solve.
//
// mfista_nufft_solve_device_noalloc runs the same solve as
// mfista_nufft_solve_device (set-up, persistent cooperative launch, FP64
// RESULT pass) but takes every workspace from `scratch`, a file-scope device
// array the code object brings with it, so the call makes no hipMalloc or
// hipFree, creates no stream or event and allocates no pinned memory: for
// callers that cannot allocate device memory (tight serving loops, harnesses
// that forbid it).  Each device that loads the library holds its own copy of
// the array until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address is looked up once per device with hipGetSymbolAddress (with
// that device current) and only that address is used.  Every call on a device
// shares that device's array and leaves the last call's contents in it; calls
// are serialised by the library mutex and each returns when its solve has
// finished, and everything that must start at zero is cleared by the solve
// itself.  Calls that overlapped on one device (e.g. from two processes
// sharing one copy of the library state) would race.
#include <cstdio>

#include "../../include/mfista.h"
#include "solver.h"

namespace mf {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // MFISTA_NOALLOC_SCRATCH_BYTES
static_assert(kScratchBytes == MFISTA_NOALLOC_SCRATCH_BYTES, "keep the header's limit in step");
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static void* g_scratch[kMaxDevices];  // device address of `scratch`, per device

// The current device's copy of `scratch` (library mutex held); nullptr and an
// error message if it cannot be found.
static void* scratch_ptr(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fprintf(stderr, "[mfista_amd] hipGetSymbolAddress(scratch) failed on device %d: %s\n", dev,
              hipGetErrorName(e));
      return nullptr;
    }
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}

}  // namespace mf

using namespace mf;

extern "C" int mfista_nufft_solve_device_noalloc(const double* d_u, const double* d_v, const double* d_vis_r,
                                                 const double* d_vis_i, const double* d_vis_std, int M, int NX,
                                                 int NY, int maxiter, double eps, double lambda_l1,
                                                 double lambda_tv, double lambda_tsv, double cinit,
                                                 const double* d_xinit, double* d_xout, int nonneg_flag,
                                                 int box_flag, const float* d_cl_box, struct RESULT* res,
                                                 mfista_opts* opts, void* stream) {
  std::lock_guard<std::mutex> lock(global_mutex());
  if (!d_u || !d_v || !d_vis_r || !d_vis_i || !d_vis_std || !d_xinit || !d_xout || !res) return MFISTA_E_ARG;
  if (box_flag && !d_cl_box) return MFISTA_E_ARG;
  if (opts && opts->engine != MFISTA_ENGINE_PERSISTENT) return MFISTA_E_ARG;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return MFISTA_E_HIP;
  void* base = scratch_ptr(dev);
  if (!base) return MFISTA_E_HIP;
  Arena arena;
  arena.base = (char*)base;
  arena.cap = kScratchBytes;
  Context C;  // host-side bookkeeping only; every DevMem of it is carved from the arena
  C.device = dev;
  C.stream = (hipStream_t)stream;
  C.no_alloc = true;
  SolveArgs A;
  A.u = d_u;
  A.v = d_v;
  A.vr = d_vis_r;
  A.vi = d_vis_i;
  A.sd = d_vis_std;
  A.M = M;
  A.NX = NX;
  A.NY = NY;
  A.maxiter = maxiter;
  A.eps = eps;
  A.lambda_l1 = lambda_l1;
  A.lambda_tv = lambda_tv;
  A.lambda_tsv = lambda_tsv;
  A.cinit = cinit;
  A.xinit = d_xinit;
  A.xout = d_xout;
  A.nonneg = nonneg_flag;
  A.box_flag = box_flag;
  A.box = d_cl_box;
  int rc;
  {
    ArenaScope scope(&arena);
    rc = solve_nufft(C, A, res, opts);
  }
  if (arena.exhausted) {
    fprintf(stderr,
            "[mfista_amd] mfista_nufft_solve_device_noalloc: the problem (%dx%d, M=%d) needs more than the "
            "%zu MiB of scratch memory\n",
            NX, NY, M, kScratchBytes >> 20);
    return MFISTA_E_SIZE;
  }
  return rc;
}

extern "C" long mfista_device_allocations(void) {
  std::lock_guard<std::mutex> lock(global_mutex());
  return device_allocations();
}
