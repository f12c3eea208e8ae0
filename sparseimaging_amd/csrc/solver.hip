// solver.hip -- host orchestration and the C ABI of libmfista_amd.so.
//
// Entry points replace c/mfista.h:151-178 and c++/mfista.hpp:117-140.  A
// per-process context owns one HIP stream and grow-only device workspaces,
// so repeated solves of the same shape allocate nothing.  The NUFFT set-up
// (reference preNUFFT, c/mfista_nufft_lib.c:50-121) runs on the GPU: cell
// indices and tap slopes per visibility, a stable sort by gridding tile for
// locality, and the per-tile visibility lists of the LDS gridding kernel.
#include <algorithm>
#include <chrono>
#include <type_traits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <set>
#include <vector>

#include "../../include/mfista.h"
#include "autotune.h"
#include "mfista_kernels.h"
#include "nufft.h"
#include "driver.h"
#include "solver.h"

namespace mf {

int sort_pairs(void* temp, size_t* temp_bytes, const int* kin, int* kout, const int* vin, int* vout,
               int n, int end_bit, hipStream_t st);
int sort_pairs_u(void* temp, size_t* temp_bytes, const int* kin, int* kout, const unsigned* vin,
                 unsigned* vout, int n, int end_bit, hipStream_t st);

// Every live workspace allocation, so that mfista_amd_finalize can free them
// all (DevMem objects live in contexts, in vectors that move them, and in
// file-scope state).  Guarded by the library mutex, which every entry point
// holds while it allocates.
static std::set<void*> g_live;
static long g_nalloc = 0;
static Arena* g_arena = nullptr;

ArenaScope::ArenaScope(Arena* a) { g_arena = a; }
ArenaScope::~ArenaScope() { g_arena = nullptr; }

long device_allocations() { return g_nalloc; }

int DevMem::ensure(size_t bytes) {
  if (bytes == 0) bytes = 16;
  if (bytes <= n) return 0;
  release();
  if (g_arena) {
    Arena& a = *g_arena;
    const size_t off = (a.off + 255) & ~(size_t)255;
    if (off > a.cap || bytes > a.cap - off) {
      a.exhausted = true;
      return -1;
    }
    p = a.base + off;
    n = bytes;
    in_arena = true;
    a.off = off + bytes;
    if (a.off > a.peak) a.peak = a.off;
    return 0;
  }
  if (hipMalloc(&p, bytes) != hipSuccess) {
    p = nullptr;
    return -1;
  }
  g_nalloc++;
  g_live.insert(p);
  n = bytes;
  return 0;
}

void DevMem::release() {
  if (p && !in_arena) {
    g_live.erase(p);
    (void)hipFree(p);
  }
  p = nullptr;
  n = 0;
  in_arena = false;
}

int upload_twiddles_pub(const FFTPlan& P, DevMem& tf, DevMem& td, hipStream_t st) {
  const int L = P.L;
  std::vector<float2> hf(L);
  std::vector<double2> hd(L);
  for (int t = 0; t < L; t++) {
    const double a = -2.0 * M_PI * (double)t / (double)L;
    hd[t] = make_double2(cos(a), sin(a));
    hf[t] = make_float2((float)cos(a), (float)sin(a));
  }
  if (tf.ensure(L * sizeof(float2)) || td.ensure(L * sizeof(double2))) return -1;
  MF_CHECK(hipMemcpyAsync(tf.p, hf.data(), L * sizeof(float2), hipMemcpyHostToDevice, st));
  MF_CHECK(hipMemcpyAsync(td.p, hd.data(), L * sizeof(double2), hipMemcpyHostToDevice, st));
  MF_CHECK(hipStreamSynchronize(st));
  return 0;
}

static int bits_for(int v) {
  int b = 1;
  while ((1 << b) <= v) b++;
  return b;
}

// ---------------------------------------------------------------------------
// NUFFT set-up on the device.  vr/vi/sd may be null (operator-only use).
int nufft_setup(Context& C, int preset, const double* d_u, const double* d_v, const double* d_vr,
                const double* d_vi, const double* d_sd, int M, int NX, int NY) {
  hipStream_t st = C.stream;
  // even sizes; 2N-point LDS transforms need 2 * 2N * 16 B <= 64 KiB in FP64
  if (M <= 0 || NX < 4 || NY < 4 || (NX % 2) || (NY % 2) || NX > 1024 || NY > 1024)
    return MFISTA_E_SIZE;
  const int MSP = (preset == MFISTA_PRESET_CXX12) ? 6 : 12;
  NufftState& N = C.nu;
  N.preset = preset;
  N.NX = NX;
  N.NY = NY;
  N.M = M;
  N.MSP = MSP;
  GridGeom& g = N.geom;
  g.NX = NX;
  g.NY = NY;
  g.M = M;
  g.MSP = MSP;
  g.Lx = 2 * NX;
  g.Ly = 2 * NY;
  // gridding tiles: 16 x 16 cells by default, one per thread of the
  // persistent gatherer; the autotuner may choose 32 x 16 or 32 x 32 (2 or 4
  // cells per thread).  Per visibility the gather evaluates (tiles touched) x
  // (cells per tile): ~730 cells for the 12-tap footprint and ~1500 for 24
  // taps with 16 x 16 tiles, 3000 for 24 taps with 32 x 32 tiles; larger tiles
  // have fewer shared edges to flush with atomics.
  g.TX = C.grid_tile_x;
  g.TY = C.grid_tile_y;
  g.ntx = (g.Lx + g.TX - 1) / g.TX;
  g.nty = (NY + 1 + g.TY - 1) / g.TY;
  const double pi = M_PI, Mrx = 2.0 * NX, Mry = 2.0 * NY;
  double taux, tauy, coeff;
  if (preset == MFISTA_PRESET_C24) {
    taux = 12.0 / ((double)NX * NX);
    tauy = 12.0 / ((double)NY * NY);
    coeff = pi / sqrt(taux * tauy);
  } else {
    taux = (double)MSP / ((double)NX * NX);
    tauy = taux;
    coeff = ((double)NX * NY) * pi / ((double)MSP);
  }
  g.t3x = (pi / Mrx) * (pi / Mrx) / taux;
  g.t3y = (pi / Mry) * (pi / Mry) / tauy;
  g.inv_MM = 1.0 / (4.0 * (double)NX * NY);
  if (!make_fft_plan(g.Lx, &N.Px) || !make_fft_plan(g.Ly, &N.Py)) return MFISTA_E_SIZE;
  if (upload_twiddles_pub(N.Px, N.twxf, N.twxd, st) || upload_twiddles_pub(N.Py, N.twyf, N.twyd, st))
    return MFISTA_E_HIP;
  {  // separable deconvolution E4(i,j) = e4x[i] * e4y[j]
    std::vector<double> ex(NX), ey(NY);
    std::vector<float> exf(NX), eyf(NY);
    for (int i = 0; i < NX; i++) {
      const double ti = (double)(i - NX / 2);
      ex[i] = coeff * exp(taux * ti * ti);
      exf[i] = (float)ex[i];
    }
    // the C library centres the second axis with Nx/2 (c/mfista_nufft_lib.c:82)
    const int cy = (preset == MFISTA_PRESET_C24) ? NX / 2 : NY / 2;
    for (int j = 0; j < NY; j++) {
      const double tj = (double)(j - cy);
      ey[j] = exp(tauy * tj * tj);
      eyf[j] = (float)ey[j];
    }
    if (N.e4xf.ensure(NX * 4) || N.e4yf.ensure(NY * 4) || N.e4xd.ensure(NX * 8) ||
        N.e4yd.ensure(NY * 8))
      return MFISTA_E_HIP;
    MF_CHECK(hipMemcpyAsync(N.e4xf.p, exf.data(), NX * 4, hipMemcpyHostToDevice, st));
    MF_CHECK(hipMemcpyAsync(N.e4yf.p, eyf.data(), NY * 4, hipMemcpyHostToDevice, st));
    MF_CHECK(hipMemcpyAsync(N.e4xd.p, ex.data(), NX * 8, hipMemcpyHostToDevice, st));
    MF_CHECK(hipMemcpyAsync(N.e4yd.p, ey.data(), NY * 8, hipMemcpyHostToDevice, st));
  }
  const size_t Ms = (size_t)M;
  const int ntiles = g.ntx * g.nty;
  const int slots = tile_slots(MSP, g.TX, g.TY);
  const size_t nent = (size_t)slots * Ms;
  if (N.mxy.ensure(Ms * sizeof(int2)) || N.af.ensure(Ms * sizeof(float2)) ||
      N.e1f.ensure(Ms * 4) || N.visf.ensure(Ms * sizeof(float2)) || N.wf.ensure(Ms * 4) ||
      N.ad.ensure(Ms * sizeof(double2)) || N.e1d.ensure(Ms * 8) ||
      N.visd.ensure(Ms * sizeof(double2)) || N.wd.ensure(Ms * 8) || N.perm.ensure(Ms * 4) ||
      N.yneg.ensure(Ms * 4) || N.keys_a.ensure(nent * 4) || N.keys_b.ensure(nent * 4) ||
      N.idx_a.ensure(Ms * 4) || N.vals_a.ensure(nent * 4) || N.ent.ensure(nent * 4) ||
      N.tstart.ensure((size_t)ntiles * 4) || N.tend.ensure((size_t)ntiles * 4) ||
      N.err.ensure(16) || N.zero_vis.ensure(Ms * sizeof(double2)))
    return MFISTA_E_HIP;
  const size_t nky = (size_t)NY + 1;
  if (N.T.ensure(nky * NX * sizeof(float2)) ||
      N.Gf.ensure(((size_t)NY + MSP) * g.Lx * sizeof(float2)) ||
      N.Hg.ensure(nky * g.Lx * sizeof(float2)) || N.U.ensure((size_t)NX * nky * sizeof(float2)) ||
      N.rw.ensure(Ms * sizeof(float2)))
    return MFISTA_E_HIP;
  N.have64 = false;
  int* err = (int*)N.err.p;
  MF_CHECK(hipMemsetAsync(err, 0, 16, st));
  launch_vis_keys(d_u, d_v, M, NX, NY, preset, MSP, g.TX, g.TY, g.ntx, (int*)N.keys_a.p,
                  (int*)N.idx_a.p, err, st);
  const int kbits = bits_for(ntiles + 1);
  size_t b1 = 0, b2 = 0;
  sort_pairs(nullptr, &b1, (int*)N.keys_a.p, (int*)N.keys_b.p, (int*)N.idx_a.p, (int*)N.perm.p,
             M, kbits, st);
  sort_pairs_u(nullptr, &b2, (int*)N.keys_a.p, (int*)N.keys_b.p, (unsigned*)N.vals_a.p,
               (unsigned*)N.ent.p, (int)nent, kbits, st);
  size_t tb = b1 > b2 ? b1 : b2;
  if (N.sort_tmp.ensure(tb)) return MFISTA_E_HIP;
  tb = N.sort_tmp.n;
  if (sort_pairs(N.sort_tmp.p, &tb, (int*)N.keys_a.p, (int*)N.keys_b.p, (int*)N.idx_a.p,
                 (int*)N.perm.p, M, kbits, st))
    return MFISTA_E_HIP;
  // visibility data: missing arrays are read as zeros / unit sigma
  const double* vr = d_vr;
  const double* vi = d_vi;
  const double* sd = d_sd;
  if (!vr || !vi || !sd) {
    if (N.ones.ensure(Ms * 8)) return MFISTA_E_HIP;
    std::vector<double> one(Ms, 1.0);
    MF_CHECK(hipMemcpyAsync(N.ones.p, one.data(), Ms * 8, hipMemcpyHostToDevice, st));
    MF_CHECK(hipMemsetAsync(N.zero_vis.p, 0, Ms * sizeof(double2), st));
    vr = (const double*)N.zero_vis.p;
    vi = (const double*)N.zero_vis.p;
    sd = (const double*)N.ones.p;
  }
  launch_vis_params((const int*)N.perm.p, d_u, d_v, vr, vi, sd, M, NX, NY, preset, MSP, N.pf(),
                    N.pd(), (int*)N.yneg.p, st);
  launch_tile_entries((const int2*)N.mxy.p, M, NX, NY, MSP, g.TX, g.TY, g.ntx, g.nty, slots,
                      (int*)N.keys_a.p, (unsigned*)N.vals_a.p, st);
  tb = N.sort_tmp.n;
  if (sort_pairs_u(N.sort_tmp.p, &tb, (int*)N.keys_a.p, (int*)N.keys_b.p, (unsigned*)N.vals_a.p,
                   (unsigned*)N.ent.p, (int)nent, kbits, st))
    return MFISTA_E_HIP;
  MF_CHECK(hipMemsetAsync(N.tstart.p, 0, (size_t)ntiles * 4, st));
  MF_CHECK(hipMemsetAsync(N.tend.p, 0, (size_t)ntiles * 4, st));
  launch_tile_ranges((const int*)N.keys_b.p, (int)nent, ntiles, (int*)N.tstart.p,
                     (int*)N.tend.p, st);
  int herr = 0;
  std::vector<int> hs(ntiles), he(ntiles);
  MF_CHECK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
  MF_CHECK(hipMemcpyAsync(hs.data(), N.tstart.p, (size_t)ntiles * 4, hipMemcpyDeviceToHost, st));
  MF_CHECK(hipMemcpyAsync(he.data(), N.tend.p, (size_t)ntiles * 4, hipMemcpyDeviceToHost, st));
  MF_CHECK(hipStreamSynchronize(st));
  MF_CHECK(hipGetLastError());
  if (herr) return MFISTA_E_FOV;
  // gridding work list: (tile, first entry, end entry) chunks of <= kGridChunk entries
  // (w: bit 0 = first chunk of its tile, bit 1 = last chunk of its tile)
  std::vector<int4>& ch = N.h_chunks;
  ch.clear();
  for (int t = 0; t < ntiles; t++)
    for (int e = hs[t]; e < he[t]; e += kGridChunk) {
      const int end = std::min(he[t], e + kGridChunk);
      ch.push_back(make_int4(t, e, end, (e == hs[t] ? 1 : 0) | (end == he[t] ? 2 : 0)));
    }
  N.nchunks = (int)ch.size();
  if (N.chunks.ensure(ch.size() * sizeof(int4) + 16)) return MFISTA_E_HIP;
  if (!ch.empty())
    MF_CHECK(hipMemcpyAsync(N.chunks.p, ch.data(), ch.size() * sizeof(int4), hipMemcpyHostToDevice, st));
  MF_CHECK(hipMemsetAsync(N.Hg.p, 0, nky * g.Lx * sizeof(float2), st));
  MF_CHECK(hipStreamSynchronize(st));
  N.ready = true;
  return 0;
}

int nufft_ensure64(Context& C) {
  NufftState& N = C.nu;
  if (N.have64) return 0;
  const GridGeom& g = N.geom;
  const size_t nky = (size_t)N.NY + 1;
  if (N.T64.ensure(nky * N.NX * sizeof(double2)) ||
      N.Gf64.ensure(((size_t)N.NY + N.MSP) * g.Lx * sizeof(double2)) ||
      N.Hg64.ensure(nky * g.Lx * sizeof(double2)) ||
      N.U64.ensure((size_t)N.NX * nky * sizeof(double2)) ||
      N.rw64.ensure((size_t)N.M * sizeof(double2)))
    return MFISTA_E_HIP;
  MF_CHECK(hipMemsetAsync(N.Hg64.p, 0, nky * g.Lx * sizeof(double2), C.stream));
  N.have64 = true;
  return 0;
}

template <typename R>
struct Sel;
template <>
struct Sel<float> {
  static VisParams<float> vp(NufftState& N) { return N.pf(); }
  static const float2* tx(NufftState& N) { return (const float2*)N.twxf.p; }
  static const float2* ty(NufftState& N) { return (const float2*)N.twyf.p; }
  static const float* ex(NufftState& N) { return (const float*)N.e4xf.p; }
  static const float* ey(NufftState& N) { return (const float*)N.e4yf.p; }
  static float2* T(NufftState& N) { return (float2*)N.T.p; }
  static float2* Gf(NufftState& N) { return (float2*)N.Gf.p; }
  static float2* Hg(NufftState& N) { return (float2*)N.Hg.p; }
  static float2* U(NufftState& N) { return (float2*)N.U.p; }
};
template <>
struct Sel<double> {
  static VisParams<double> vp(NufftState& N) { return N.pd(); }
  static const double2* tx(NufftState& N) { return (const double2*)N.twxd.p; }
  static const double2* ty(NufftState& N) { return (const double2*)N.twyd.p; }
  static const double* ex(NufftState& N) { return (const double*)N.e4xd.p; }
  static const double* ey(NufftState& N) { return (const double*)N.e4yd.p; }
  static double2* T(NufftState& N) { return (double2*)N.T64.p; }
  static double2* Gf(NufftState& N) { return (double2*)N.Gf64.p; }
  static double2* Hg(NufftState& N) { return (double2*)N.Hg64.p; }
  static double2* U(NufftState& N) { return (double2*)N.U64.p; }
};

// A x: residual rz = w (V' - A x) in R precision and/or its FP32 copy rz32
// (all optional), chi^2 partials (optional), model visibilities in input order
// (optional)
template <typename R>
void nufft_fwd(Context& C, const R* x, cplx<R>* rz, double* chi_part, cplx<R>* model,
               const int* skip, float2* rz32) {
  NufftState& N = C.nu;
  const GridGeom& g = N.geom;
  launch_rowfft_fwd<R>(x, Sel<R>::ex(N), Sel<R>::ey(N), Sel<R>::T(N), N.NX, N.NY, g.Ly, N.NY / 2,
                       N.Py, Sel<R>::ty(N), C.stream, skip);
  launch_colfft_fwd<R>(Sel<R>::T(N), Sel<R>::Gf(N), N.NX, N.NY + 1, g.Lx, N.NX / 2, 0, N.MSP,
                       N.Px, Sel<R>::tx(N), C.stream, skip);
  launch_degrid<R>(g, Sel<R>::vp(N), Sel<R>::Gf(N), rz, rz32, nullptr, chi_part, nullptr, 0,
                   degrid_blocks(N.M, N.MSP), model, (const int*)N.perm.p, (const int*)N.yneg.p,
                   C.stream, skip);
}

// A d for a step d = xnew - z (FP32): partials of |rz64 - w A d|^2 (chi^2 at
// xnew, with rz64 the FP64 residual at z) and of |w A d|^2
template <typename R>
void nufft_fwd_delta(Context& C, const R* d, double* chi_part, double* lip_part, const int* skip) {
  NufftState& N = C.nu;
  const GridGeom& g = N.geom;
  launch_rowfft_fwd<R>(d, Sel<R>::ex(N), Sel<R>::ey(N), Sel<R>::T(N), N.NX, N.NY, g.Ly, N.NY / 2,
                       N.Py, Sel<R>::ty(N), C.stream, skip);
  launch_colfft_fwd<R>(Sel<R>::T(N), Sel<R>::Gf(N), N.NX, N.NY + 1, g.Lx, N.NX / 2, 0, N.MSP,
                       N.Px, Sel<R>::tx(N), C.stream, skip);
  launch_degrid<R>(g, Sel<R>::vp(N), Sel<R>::Gf(N), nullptr, nullptr, (const double2*)N.rw64.p,
                   chi_part, lip_part, 1, degrid_blocks(N.M, N.MSP), nullptr, nullptr, nullptr,
                   C.stream, skip);
}

// g = Re(A^H w r) from the reflected-frame residual rz
template <typename R>
void nufft_adj(Context& C, const cplx<R>* rz, R* gout, const int* skip) {
  NufftState& N = C.nu;
  const GridGeom& g = N.geom;
  launch_grid<R>(g, Sel<R>::vp(N), rz, (const int4*)N.chunks.p, N.nchunks, (const unsigned*)N.ent.p,
                 Sel<R>::Hg(N), C.stream, skip);
  launch_colfft_inv<R>(Sel<R>::Hg(N), Sel<R>::U(N), N.NX, N.NY + 1, g.Lx, N.NX / 2, 0, N.Px,
                       Sel<R>::tx(N), C.stream, skip);
  launch_rowfft_c2r<R>(Sel<R>::U(N), Sel<R>::ex(N), Sel<R>::ey(N),
                       (R)(1.0 / (8.0 * (double)N.NX * N.NY)), gout, N.NX, N.NY, g.Ly, N.NY / 2,
                       N.Py, Sel<R>::ty(N), C.stream, skip);
}

template void nufft_fwd<float>(Context&, const float*, float2*, double*, float2*, const int*, float2*);
template void nufft_fwd<double>(Context&, const double*, double2*, double*, double2*, const int*, float2*);
template void nufft_adj<float>(Context&, const float2*, float*, const int*);
template void nufft_fwd_delta<float>(Context&, const float*, double*, double*, const int*);
template void nufft_fwd_delta<double>(Context&, const double*, double*, double*, const int*);
template void nufft_adj<double>(Context&, const double2*, double*, const int*);

// gather y (input order) into the reflected frame, sorted order
template <typename R>
__global__ void k_gather_vis(const double* __restrict__ yr, const double* __restrict__ yi,
                             const int* __restrict__ perm, const int* __restrict__ yneg, int M,
                             cplx<R>* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  const int k = perm[s];
  const double im = yneg[s] ? -yi[k] : yi[k];
  out[s] = mk<R>((R)yr[k], (R)im);
}

template <typename R>
__global__ void k_split_model(const cplx<R>* __restrict__ m, double* __restrict__ yr,
                              double* __restrict__ yi, int M) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  yr[k] = (double)m[k].x;
  yi[k] = (double)m[k].y;
}

// ---------------------------------------------------------------------------
static double host_sum(const std::vector<double>& v, int off, int cnt, int stride) {
  double s = 0;
  for (int i = 0; i < cnt; i++) s += v[(size_t)off + (size_t)i * stride];
  return s;
}

// FP64 RESULT fields from a device FP64 image (the reference's calc_result_*).
// chi2 = sum_k w_k^2 |V_k - (A x)_k|^2 is supplied by the caller.
static void fill_result_common(RESULT* R, int NX, int NY, double lambda_l1, double lambda_tv,
                               double lambda_tsv, double chi2, const double stats[4]) {
  R->N = NX * NY;
  R->NX = NX;
  R->NY = NY;
  R->lambda_l1 = lambda_l1;
  R->lambda_tv = lambda_tv;
  R->lambda_tsv = lambda_tsv;
  R->sq_error = chi2;
  R->l1cost = stats[0];
  R->N_active = (int)llround(stats[3]);
  R->finalcost = chi2 / 2;
  if (lambda_l1 > 0) R->finalcost += lambda_l1 * R->l1cost;
  if (lambda_tsv > 0) {
    R->tsvcost = stats[1];
    R->finalcost += lambda_tsv * R->tsvcost;
  } else if (lambda_tv > 0) {
    R->tvcost = stats[2];
    R->finalcost += lambda_tv * R->tvcost;
  }
}

int image_stats64(Context& C, const double* d_x64, int NX, int NY, double out[4]) {
  const int nb = elem_blocks(NX * NY);
  if (C.stat64.ensure((size_t)4 * nb * 8)) return MFISTA_E_HIP;
  launch_stats<double>(d_x64, NX, NY, (double*)C.stat64.p, C.stream);
  std::vector<double> h((size_t)4 * nb);
  MF_CHECK(hipMemcpyAsync(h.data(), C.stat64.p, h.size() * 8, hipMemcpyDeviceToHost, C.stream));
  MF_CHECK(hipStreamSynchronize(C.stream));
  for (int q = 0; q < 4; q++) out[q] = host_sum(h, q, nb, 4);
  return 0;
}

int nufft_chi2_64(Context& C, const double* d_x64, double* chi2) {
  NufftState& N = C.nu;
  if (nufft_ensure64(C)) return MFISTA_E_HIP;
  const int nb = degrid_blocks(N.M, N.MSP);
  if (C.chi64.ensure((size_t)nb * 8)) return MFISTA_E_HIP;
  nufft_fwd<double>(C, d_x64, nullptr, (double*)C.chi64.p, nullptr, nullptr);
  std::vector<double> h(nb);
  MF_CHECK(hipMemcpyAsync(h.data(), C.chi64.p, nb * 8, hipMemcpyDeviceToHost, C.stream));
  MF_CHECK(hipStreamSynchronize(C.stream));
  *chi2 = host_sum(h, 0, nb, 1);
  return 0;
}

// ---------------------------------------------------------------------------
// The momentum-point residual (the objective of the backtracking test and the
// input of the gradient) is evaluated with an FP64 forward NUFFT; the step
// transform, the gradient and the prox stay FP32.
// The stream engine's NUFFT operator.  Iterates are FP64 (FP32 iterates
// stagnate near a tight tolerance: the prox step falls below their rounding);
// the residual at the momentum point comes from an FP64 forward transform,
// the gradient from the FP32 adjoint of its FP32 copy, and the trial's A d
// from an FP64 forward transform of the step.
struct NufftOp : Operator<double> {
  Context& C;
  explicit NufftOp(Context& c) : C(c) {}
  int npart() const override { return degrid_blocks(C.nu.M, C.nu.MSP); }
  void fwd(const double* x, bool keep, double* part, const int* skip) override {
    nufft_fwd<double>(C, x, keep ? (double2*)C.nu.rw64.p : nullptr, part, nullptr, skip,
                      keep ? (float2*)C.nu.rw.p : nullptr);
  }
  void adj(double* g, const int* skip) override {
    float* g32 = (float*)C.g32.p;
    nufft_adj<float>(C, (const float2*)C.nu.rw.p, g32, skip);
    launch_cast<float, double>(g32, g, C.nu.NX * C.nu.NY, C.stream, skip);
  }
  bool has_delta() const override { return true; }
  void fwd_delta(const double* d, double* chi_part, double* lip_part, const int* skip) override {
    nufft_fwd_delta<double>(C, d, chi_part, lip_part, skip);
  }
};

// Whether a single persistent solve may run the Toeplitz loop (opts->loop
// permitting): L1 and L1+TSV, paired rounds, floating-point atomics allowed in
// its set-up, square images up to 512^2 (the presets size both axes' kernels
// from NX, so for NX != NY the reference's transform is not the exact one the
// loop applies) of even size (the pruned transforms place pixel i at
// (i - N/2) mod 2N with N/2 exact), and not the allocation-free entry point.
static bool toeplitz_applies(const Context& C, const SolveArgs& A, const mfista_opts* opts) {
  const bool tv = A.lambda_tv != 0;
  return !tv && !C.no_alloc && !(opts && (opts->single_trials || opts->deterministic)) &&
         (!opts || opts->loop != MFISTA_LOOP_GRIDDING) && A.NX == A.NY && A.NX <= 512 && A.NX % 2 == 0;
}

// Times the autotuner's candidates on this problem: `iters` iterations from
// its xinit, no stopping rule, xout overwritten.
static int tune_single(Context& C, const SolveArgs& A, const mfista_opts* opts, int iters, TuneChoice* best,
                       std::vector<TuneChoice>* all) {
  SolveArgs a = A;
  a.maxiter = iters > 0 ? iters : 30;
  a.eps = 0.0;
  mfista_opts o;
  if (opts)
    o = *opts;
  else
    mfista_default_opts(&o);
  o.autotune = 0;
  o.rel_tol = 0.0;
  o.engine = MFISTA_ENGINE_PERSISTENT;
  o.verbose = 0;
  auto run = [&](double* us) -> int {
    RESULT r;
    memset(&r, 0, sizeof(r));
    const int rc = solve_nufft(C, a, &r, &o);
    if (rc) return rc;
    if (o.engine_used != MFISTA_ENGINE_PERSISTENT) return MFISTA_E_COOP;
    *us = o.kernel_ms * 1e3 / (r.ITER > 0 ? r.ITER : 1);
    return 0;
  };
  return tune_search(C, tune_reg(A.lambda_tv, A.lambda_tsv), false, run, best, all, toeplitz_applies(C, A, opts));
}

int solve_nufft(Context& C, const SolveArgs& A, RESULT* res, mfista_opts* opts) {
  const int preset = opts ? opts->preset : MFISTA_PRESET_C24;
  int tv_mode;
  if (A.lambda_tv == 0)
    tv_mode = 0;
  else if (A.lambda_tsv == 0)
    tv_mode = 1;
  else {
    printf("You cannot set both of lambda_TV and lambda_TSV positive.\n");
    return MFISTA_E_BOTH_REG;
  }
  if (tv_mode && preset == MFISTA_PRESET_CXX12) {
    printf("We have not implemented TV option.\n");
    return MFISTA_E_TV_PRESET;
  }
  if (A.maxiter <= 0) return MFISTA_E_ARG;
  // the autotuner's launch configuration for this shape (tuned here on a
  // miss; the allocation-free entry point only looks it up)
  const int engine_req = opts ? opts->engine : MFISTA_ENGINE_PERSISTENT;
  bool tuned = false;
  if (opts && opts->autotune && engine_req == MFISTA_ENGINE_PERSISTENT) {
    const TuneKey k = tune_key(A.NX, A.NY, preset, tune_reg(A.lambda_tv, A.lambda_tsv), A.M, 1);
    TuneChoice ch;
    bool have = tune_lookup(k, &ch);
    if (!have && !C.no_alloc) {
      const int trc = tune_single(C, A, opts, opts->tune_iters, &ch, nullptr);
      if (trc) return trc;
      tune_store(k, ch);
      have = true;
    }
    if (have) {
      tune_apply(C, ch);
      tuned = true;
    }
  }
  struct RestoreKnobs {
    Context& C;
    bool on;
    ~RestoreKnobs() {
      if (on) tune_restore(C);
    }
  } restore_knobs{C, tuned};
  auto t0 = std::chrono::steady_clock::now();
  int rc = nufft_setup(C, preset, A.u, A.v, A.vr, A.vi, A.sd, A.M, A.NX, A.NY);
  if (rc) {
    if (rc == MFISTA_E_FOV) printf("Error: the field of view is too small for the (u,v) coverage.\n");
    return rc;
  }
  // eps scaled by the mean |V|^2 (c/mfista_nufft_lib.c:923-926); |W V|^2 for
  // the Toeplitz loop
  SolveArgs B = A;
  double wv2 = 0;
  {
    double s = 0;
    std::vector<double> hr(A.M), hi(A.M), hs(A.M);
    MF_CHECK(hipMemcpyAsync(hr.data(), A.vr, A.M * 8, hipMemcpyDeviceToHost, C.stream));
    MF_CHECK(hipMemcpyAsync(hi.data(), A.vi, A.M * 8, hipMemcpyDeviceToHost, C.stream));
    MF_CHECK(hipMemcpyAsync(hs.data(), A.sd, A.M * 8, hipMemcpyDeviceToHost, C.stream));
    MF_CHECK(hipStreamSynchronize(C.stream));
    for (int i = 0; i < A.M; i++) s += hr[i] * hr[i];
    for (int i = 0; i < A.M; i++) s += hi[i] * hi[i];
    B.eps_abs = A.eps * s / (double)A.M;
    for (int i = 0; i < A.M; i++) wv2 += (hr[i] * hr[i] + hi[i] * hi[i]) / (hs[i] * hs[i]);
  }
  RunStats rs;
  rs.c = A.cinit;
  if (!tv_mode && A.nonneg != 0 && A.nonneg != 1) {
    printf("nonneg_flag must be chosen properly.\n");
    rc = 0;
  } else {
    if (nufft_ensure64(C)) return MFISTA_E_HIP;
    // the stream engine's buffers (the allocation-free entry point has no stream engine)
    if (!C.no_alloc && (C.x64.ensure((size_t)A.NX * A.NY * 8) || C.g32.ensure((size_t)A.NX * A.NY * 4)))
      return MFISTA_E_HIP;
    NufftOp op(C);
    RunParams P;
    P.NX = A.NX;
    P.NY = A.NY;
    P.maxiter = A.maxiter;
    P.nonneg = A.nonneg;
    P.box_flag = A.box_flag;
    P.tv_mode = tv_mode;
    P.fgp_iter = (opts && opts->fgp_iter > 0) ? opts->fgp_iter : FGPITER;
    P.eps_abs = B.eps_abs;
    P.rel_tol = opts ? opts->rel_tol : 0.0;
    P.lambda_l1 = A.lambda_l1;
    P.lambda_tv = A.lambda_tv;
    P.lambda_tsv = A.lambda_tsv;
    P.cinit = A.cinit;
    P.eta_first = (preset == MFISTA_PRESET_CXX12) ? 10.0 : kEta;
    P.single_trials = (opts && opts->single_trials) ? 1 : 0;
    P.deterministic = (opts && opts->deterministic) ? 1 : 0;
    P.fgp_tw = C.fgp_tw;
    P.xinit = A.xinit;
    P.xout = A.xout;
    P.box = A.box;
    int coop = 0, dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
    const int engine = opts ? opts->engine : MFISTA_ENGINE_PERSISTENT;
    rc = kCoopRefused;
    int used = MFISTA_ENGINE_PERSISTENT, loop_used = 0;
    if (engine == MFISTA_ENGINE_PERSISTENT && coop) {
      // the Toeplitz loop where it applies (the faster loop at every size and M
      // measured), unless the autotuner's choice for this shape is the
      // gridding loop and opts->loop does not ask for the Toeplitz loop
      const bool forced = opts && opts->loop == MFISTA_LOOP_TOEPLITZ;
      bool toep = toeplitz_applies(C, A, opts) && (forced || C.tune_loop != MFISTA_LOOP_GRIDDING);
      if (toep) {
        const int trc = toeplitz_setup(C, A.u, A.v, A.vr, A.vi, A.sd, A.M, A.NX, A.NY, preset, wv2, C.tw.out);
        if (trc == MFISTA_E_SIZE || trc == MFISTA_E_FOV)
          toep = false;  // the set-up transform does not apply: the gridding loop
        else if (trc)
          return trc;
      }
      if (toep) {
        rc = solve_nufft_toeplitz(C, P, &rs);
        loop_used = MFISTA_LOOP_TOEPLITZ;
      } else {
        rc = solve_nufft_persistent(C, P, &rs);
        loop_used = MFISTA_LOOP_GRIDDING;
      }
    }
    if (rc == kCoopRefused && !C.no_alloc) {  // the stream engine: the same loop, one launch per stage
      if (engine == MFISTA_ENGINE_PERSISTENT && coop)
        fprintf(stderr, "[mfista_amd] solving with the stream engine instead\n");
      rs = RunStats();
      used = MFISTA_ENGINE_STREAM;
      loop_used = 0;
      rc = mfista_run<double>(C, op, P, &rs);
    }
    if (opts) {
      opts->engine_used = used;
      opts->loop_used = loop_used;
    }
    C.last_steps = rs.steps;
  }
  res->ITER = rs.ITER;
  res->Lip_const = rs.c;
  if (opts) {
    opts->max_trials = rs.max_trials;
    opts->total_trials = rs.total_trials;
    opts->kernel_ms = rs.kernel_ms;
    opts->trial_rounds = rs.rounds;
  }
  if (rc) return rc;
  MF_CHECK(hipStreamSynchronize(C.stream));
  auto t1 = std::chrono::steady_clock::now();
  const double secs = std::chrono::duration<double>(t1 - t0).count();
  if (opts) opts->solve_ms = secs * 1e3;
  res->comp_time = secs;
  res->nonneg = A.nonneg;
  res->maxiter = A.maxiter;
  // FP64 result pass (calc_result_nufft, c/mfista_nufft_lib.c:791-905)
  double chi2 = 0, stats[4] = {0, 0, 0, 0};
  const double* xo = A.xout;
  if ((rc = nufft_chi2_64(C, xo, &chi2))) return rc;
  if ((rc = image_stats64(C, xo, A.NX, A.NY, stats))) return rc;
  fill_result_common(res, A.NX, A.NY, A.lambda_l1, A.lambda_tv, A.lambda_tsv, chi2, stats);
  res->M = (preset == MFISTA_PRESET_CXX12) ? A.M : (int)(A.M / 2);
  res->mean_sq_error = res->sq_error / (double)A.M;
  return 0;
}

// ---------------------------------------------------------------------------
static std::mutex g_mu;
static std::vector<Context*> g_ctx;

Context* get_context() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  if ((int)g_ctx.size() <= dev) g_ctx.resize(dev + 1, nullptr);
  if (!g_ctx[dev]) {
    Context* C = new Context();
    C->device = dev;
    if (hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking) != hipSuccess) {
      delete C;
      return nullptr;
    }
    if (hipHostMalloc((void**)&C->hflags, 4 * sizeof(int), hipHostMallocDefault) != hipSuccess) {
      delete C;
      return nullptr;
    }
    (void)hipEventCreateWithFlags(&C->ev[0], hipEventDisableTiming);
    (void)hipEventCreateWithFlags(&C->ev[1], hipEventDisableTiming);
    // tuning knob: workgroups per CU of the persistent solver (1 to 4, capped by occupancy)
    if (const char* e = getenv("MFISTA_PERSISTENT_BLOCKS_PER_CU")) {
      const int v = atoi(e);
      if (v >= 1 && v <= 4) C->persist_blocks_per_cu = v;
    }
    if (const char* e = getenv("MFISTA_PERSISTENT_FLAT_BARRIER")) C->persist_hier_barrier = atoi(e) ? 0 : 1;
    if (const char* e = getenv("MFISTA_PERSISTENT_PAIRED")) C->persist_pair = atoi(e) ? 1 : 0;
    C->def_blocks_per_cu = C->persist_blocks_per_cu;
    C->def_pair = C->persist_pair;
    tune_load_env();
    g_ctx[dev] = C;
  }
  return g_ctx[dev];
}

std::mutex& global_mutex() { return g_mu; }

void reset_fft_state();  // driver.hip

// Releases every context and workspace (library mutex held): waits for the
// library's work, frees device and pinned memory, destroys the streams and
// events.  The next call creates a fresh context.
static void finalize_all() {
  for (Context*& C : g_ctx) {
    if (!C) continue;
    (void)hipSetDevice(C->device);
    if (C->stream) (void)hipStreamSynchronize(C->stream);
    for (hipEvent_t& e : C->ev)
      if (e) (void)hipEventDestroy(e);
    for (hipEvent_t& e : C->ev_t)
      if (e) (void)hipEventDestroy(e);
    if (C->hflags) (void)hipHostFree(C->hflags);
    if (C->stream) (void)hipStreamDestroy(C->stream);
    delete C;
    C = nullptr;
  }
  g_ctx.clear();
  for (void* p : g_live) (void)hipFree(p);
  g_live.clear();
  reset_fft_state();
}

// upload helper for the host-pointer entry points
int upload(Context& C, DevMem& m, const void* h, size_t bytes) {
  if (m.ensure(bytes)) return MFISTA_E_HIP;
  if (h) MF_CHECK(hipMemcpyAsync(m.p, h, bytes, hipMemcpyHostToDevice, C.stream));
  return 0;
}

}  // namespace mf

using namespace mf;

extern "C" {

const char* mfista_amd_version(void) { return "mfista_amd 0.2 (gfx950)"; }

void mfista_amd_finalize(void) {
  std::lock_guard<std::mutex> lock(global_mutex());
  finalize_all();
}

void mfista_default_opts(mfista_opts* o) {
  memset(o, 0, sizeof(*o));
  o->preset = MFISTA_PRESET_C24;
  o->fgp_iter = FGPITER;
  o->rel_tol = 0.0;
  o->engine = MFISTA_ENGINE_PERSISTENT;
  o->tune_iters = 30;
}

int mfista_nufft_solve_device(const double* d_u, const double* d_v, const double* d_vis_r,
                              const double* d_vis_i, const double* d_vis_std, int M, int NX,
                              int NY, int maxiter, double eps, double lambda_l1,
                              double lambda_tv, double lambda_tsv, double cinit,
                              const double* d_xinit, double* d_xout, int nonneg_flag,
                              int box_flag, const float* d_cl_box, struct RESULT* res,
                              mfista_opts* opts, void* stream) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (!d_u || !d_v || !d_vis_r || !d_vis_i || !d_vis_std || !d_xinit || !d_xout || !res)
    return MFISTA_E_ARG;
  if (box_flag && !d_cl_box) return MFISTA_E_ARG;
  // the work is queued on the caller's stream, NULL being the HIP null stream:
  // either way it is ordered after the caller's earlier work on that stream
  // (the library's own stream is a non-blocking one, unordered with respect
  // to the caller's writes of the inputs)
  hipStream_t saved = C->stream;
  C->stream = (hipStream_t)stream;
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
  int rc = solve_nufft(*C, A, res, opts);
  C->stream = saved;
  return rc;
}

// Batched NUFFT solves (one persistent launch); include/mfista.h describes
// the contract.  Each problem follows mfista_imaging_core_nufft
// (c/mfista_nufft_lib.c:909-957): eps scaled by its mean |V|^2, the same
// argument checks and messages, and the FP64 RESULT pass.
// One pass of a batch: the NUFFT set-ups (one per visibility set), each
// problem's workspace, the launch, and (results = true) the FP64 RESULT pass.
// iters > 0 runs every problem for exactly that many iterations with no
// stopping rule and writes no results (the autotuner's candidates).
// Whether a batch's persistent launch may run the Toeplitz loop (the rule of
// single solves, for every problem).
static bool toeplitz_batch_applies(int nprob, const mfista_nufft_problem* pr, const std::vector<int>& tv,
                                   const mfista_opts* opts) {
  if (opts && (opts->single_trials || opts->deterministic || opts->loop == MFISTA_LOOP_GRIDDING)) return false;
  for (int i = 0; i < nprob; i++)
    if (tv[i] || pr[i].NX != pr[i].NY || pr[i].NX > 512 || pr[i].NX % 2) return false;
  return nprob > 0;
}

static int run_batch(Context* C, int nprob, const mfista_nufft_problem* pr, const std::vector<int>& tv,
                     mfista_opts* opts, int iters, bool results, RESULT* res, double* us_per_iter) {
  const int preset = opts ? opts->preset : MFISTA_PRESET_C24;
  const int fgp_iter = (opts && opts->fgp_iter > 0) ? opts->fgp_iter : FGPITER;
  const double rel_tol = (opts && iters <= 0) ? opts->rel_tol : 0.0;
  auto t0 = std::chrono::steady_clock::now();
  // visibility sets: problems with the same arrays share one NUFFT set-up
  std::vector<int> set_of(nprob, -1);
  std::vector<int> set_rep;  // a problem of each set
  for (int i = 0; i < nprob; i++) {
    for (size_t k = 0; k < set_rep.size(); k++) {
      const mfista_nufft_problem &a = pr[i], &b = pr[set_rep[k]];
      if (a.u_dx == b.u_dx && a.v_dy == b.v_dy && a.vis_r == b.vis_r && a.vis_i == b.vis_i &&
          a.vis_std == b.vis_std && a.M == b.M && a.NX == b.NX && a.NY == b.NY) {
        set_of[i] = (int)k;
        break;
      }
    }
    if (set_of[i] < 0) {
      set_of[i] = (int)set_rep.size();
      set_rep.push_back(i);
    }
  }
  const int nset = (int)set_rep.size();
  if ((int)C->bset.size() < nset) C->bset.resize(nset);
  int rc;
  for (int k = 0; k < nset; k++) {
    const mfista_nufft_problem& q = pr[set_rep[k]];
    BatchSet& S = C->bset[k];
    const size_t Mb = (size_t)q.M * 8;
    if ((rc = upload(*C, S.u, q.u_dx, Mb)) || (rc = upload(*C, S.v, q.v_dy, Mb)) ||
        (rc = upload(*C, S.vr, q.vis_r, Mb)) || (rc = upload(*C, S.vi, q.vis_i, Mb)) ||
        (rc = upload(*C, S.sd, q.vis_std, Mb)))
      return rc;
    double pw = 0;
    for (int m = 0; m < q.M; m++) pw += q.vis_r[m] * q.vis_r[m] + q.vis_i[m] * q.vis_i[m];
    S.vis_power = pw;
    std::swap(C->nu, S.nu);
    rc = nufft_setup(*C, preset, (const double*)S.u.p, (const double*)S.v.p, (const double*)S.vr.p,
                     (const double*)S.vi.p, (const double*)S.sd.p, q.M, q.NX, q.NY);
    std::swap(C->nu, S.nu);
    if (rc) {
      if (rc == MFISTA_E_FOV) printf("Error: the field of view is too small for the (u,v) coverage.\n");
      return rc;
    }
  }
  if ((int)C->bwork.size() < nprob) C->bwork.resize(nprob);
  std::vector<NufftState*> sets;
  std::vector<RunParams> P;
  std::vector<int> which;  // problem index of each solve
  for (int i = 0; i < nprob; i++) {
    const mfista_nufft_problem& q = pr[i];
    if (results) memset(&res[i], 0, sizeof(RESULT));
    if (!tv[i] && q.nonneg_flag != 0 && q.nonneg_flag != 1) {  // c/mfista_nufft_lib.c:634-637
      printf("nonneg_flag must be chosen properly.\n");
      continue;
    }
    PersistWork& W = C->bwork[i];
    const size_t nb = (size_t)q.NX * q.NY * 8;
    if ((rc = upload(*C, W.x0, q.xinit, nb)) || (rc = upload(*C, W.xo, nullptr, nb))) return rc;
    if (q.box_flag && (rc = upload(*C, W.box, q.cl_box, (size_t)q.NX * q.NY * 4))) return rc;
    RunParams R;
    R.NX = q.NX;
    R.NY = q.NY;
    R.maxiter = iters > 0 ? iters : q.maxiter;
    R.nonneg = q.nonneg_flag;
    R.box_flag = q.box_flag;
    R.tv_mode = tv[i];
    R.fgp_iter = fgp_iter;
    R.eps_abs = iters > 0 ? 0.0 : q.eps * C->bset[set_of[i]].vis_power / (double)q.M;
    R.rel_tol = rel_tol;
    R.lambda_l1 = q.lambda_l1;
    R.lambda_tv = q.lambda_tv;
    R.lambda_tsv = q.lambda_tsv;
    R.cinit = q.cinit;
    R.eta_first = (preset == MFISTA_PRESET_CXX12) ? 10.0 : kEta;
    R.single_trials = (opts && opts->single_trials) ? 1 : 0;
    R.deterministic = (opts && opts->deterministic) ? 1 : 0;
    R.fgp_tw = C->fgp_tw;
    R.target_F = iters > 0 ? 0.0 : q.target_objective;
    R.xinit = (const double*)W.x0.p;
    R.xout = (double*)W.xo.p;
    R.box = q.box_flag ? (const float*)W.box.p : nullptr;
    sets.push_back(&C->bset[set_of[i]].nu);
    P.push_back(R);
    which.push_back(i);
  }
  std::vector<RunStats> st;
  float kms = 0;
  // Engines: the wide engine (every kernel launched once for all problems)
  // takes batches of L1 / L1+TSV problems of one image size; the persistent
  // engine runs the batch in cooperative launches.  Both run Toeplitz steps
  // when every problem is L1 / L1+TSV on a square image up to 512^2 (the rule
  // of single solves), else gridding steps.  Target objectives are a control
  // rule of the wide engine; with Toeplitz steps the objective they are
  // compared with applies the exact operator, within ~1e-8 (relative) of the
  // preset's objective that the RESULT pass reports.
  const int engine = C->tune_engine >= 0 ? C->tune_engine : (opts ? opts->engine : MFISTA_ENGINE_PERSISTENT);
  bool targets = false, wide_ok = !P.empty();
  bool toep = !P.empty() && !(opts && (opts->single_trials || opts->deterministic)) &&
              !(opts && opts->loop == MFISTA_LOOP_GRIDDING) &&
              ((opts && opts->loop == MFISTA_LOOP_TOEPLITZ) || C->tune_loop != MFISTA_LOOP_GRIDDING);
  for (size_t j = 0; j < P.size(); j++) {
    if (P[j].target_F > 0) targets = true;
    if (P[j].tv_mode || P[j].NX != P[0].NX || P[j].NY != P[0].NY) wide_ok = false;
    if (P[j].tv_mode || P[j].NX != P[j].NY || P[j].NX > 512 || P[j].NX % 2) toep = false;
  }
  if (targets && !wide_ok) {
    fprintf(stderr, "[mfista_amd] target objectives need L1 or L1+TSV problems of one image size\n");
    return MFISTA_E_ARG;
  }
  const bool wide = wide_ok && (targets || engine == MFISTA_ENGINE_STREAM);
  if (toep) {  // the operator data of every visibility set a problem uses
    std::vector<char> used(nset, 0);
    for (int i : which) used[set_of[i]] = 1;
    for (int k = 0; k < nset && toep; k++) {
      if (!used[k]) continue;
      const mfista_nufft_problem& q = pr[set_rep[k]];
      double wv2 = 0;
      for (int m = 0; m < q.M; m++)
        wv2 += (q.vis_r[m] * q.vis_r[m] + q.vis_i[m] * q.vis_i[m]) / (q.vis_std[m] * q.vis_std[m]);
      BatchSet& S = C->bset[k];
      std::swap(C->nu, S.nu);
      rc = toeplitz_setup(*C, (const double*)S.u.p, (const double*)S.v.p, (const double*)S.vr.p,
                          (const double*)S.vi.p, (const double*)S.sd.p, q.M, q.NX, q.NY, preset, wv2, S.toep);
      std::swap(C->nu, S.nu);
      if (rc == MFISTA_E_SIZE || rc == MFISTA_E_FOV)
        toep = false;  // the set-up transform does not apply: the gridding loop
      else if (rc)
        return rc;
    }
  }
  if (P.empty()) {
    st.clear();
  } else if (wide) {
    std::vector<const ToepOut*> ops;
    for (int i : which) ops.push_back(&C->bset[set_of[i]].toep);
    rc = solve_nufft_wide(*C, sets, P, st, &kms, toep ? &ops : nullptr);
  } else if (toep) {
    std::vector<const ToepOut*> ops;
    for (int i : which) ops.push_back(&C->bset[set_of[i]].toep);
    rc = solve_nufft_toeplitz_batch(*C, sets, ops, P, st, &kms);
  } else {
    rc = solve_nufft_batch(*C, sets, P, st, &kms);
  }
  if (rc) return rc;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  int total_trials = 0, max_trials = 0, rounds = 0;
  for (size_t j = 0; results && j < which.size(); j++) {
    const int i = which[j];
    const mfista_nufft_problem& q = pr[i];
    PersistWork& W = C->bwork[i];
    RESULT* R = &res[i];
    R->ITER = st[j].ITER;
    R->Lip_const = st[j].c;
    R->nonneg = q.nonneg_flag;
    R->maxiter = q.maxiter;
    R->comp_time = secs;
    total_trials += st[j].total_trials;
    rounds += st[j].rounds;
    max_trials = std::max(max_trials, st[j].max_trials);
    MF_CHECK(hipMemcpy(q.xout, W.xo.p, (size_t)q.NX * q.NY * 8, hipMemcpyDeviceToHost));
    // FP64 result pass (calc_result_nufft, c/mfista_nufft_lib.c:791-905)
    double chi2 = 0, stats[4] = {0, 0, 0, 0};
    NufftState& N = C->bset[set_of[i]].nu;
    std::swap(C->nu, N);
    rc = nufft_chi2_64(*C, (const double*)W.xo.p, &chi2);
    if (!rc) rc = image_stats64(*C, (const double*)W.xo.p, q.NX, q.NY, stats);
    std::swap(C->nu, N);
    if (rc) return rc;
    fill_result_common(R, q.NX, q.NY, q.lambda_l1, q.lambda_tv, q.lambda_tsv, chi2, stats);
    R->M = (preset == MFISTA_PRESET_CXX12) ? q.M : (int)(q.M / 2);
    R->mean_sq_error = R->sq_error / (double)q.M;
  }
  if (opts) {
    opts->engine_used = wide ? MFISTA_ENGINE_STREAM : MFISTA_ENGINE_PERSISTENT;
    opts->loop_used = toep ? MFISTA_LOOP_TOEPLITZ : MFISTA_LOOP_GRIDDING;
    opts->solve_ms = secs * 1e3;
    opts->kernel_ms = kms;
    opts->total_trials = total_trials;
    opts->max_trials = max_trials;
    opts->trial_rounds = rounds;
  }
  if (us_per_iter) *us_per_iter = kms * 1e3 / (iters > 0 ? iters : 1);
  return 0;
}

// Argument checks of a batch (each problem as mfista_imaging_core_nufft_ex
// checks it); tv[i] = 1 for the L1+TV problems.
static int batch_check(int nprob, const mfista_nufft_problem* pr, int preset, std::vector<int>* tvp) {
  std::vector<int>& tv = *tvp;
  tv.assign(nprob, 0);
  for (int i = 0; i < nprob; i++) {
    const mfista_nufft_problem& q = pr[i];
    if (q.M <= 0 || q.NX <= 0 || q.NY <= 0 || !q.u_dx || !q.v_dy || !q.vis_r || !q.vis_i || !q.vis_std ||
        !q.xinit || !q.xout || (q.box_flag && !q.cl_box) || q.maxiter <= 0)
      return MFISTA_E_ARG;
    if (q.lambda_tv == 0)
      tv[i] = 0;
    else if (q.lambda_tsv == 0)
      tv[i] = 1;
    else {
      printf("You cannot set both of lambda_TV and lambda_TSV positive.\n");
      return MFISTA_E_BOTH_REG;
    }
    if (tv[i] && preset == MFISTA_PRESET_CXX12) {
      printf("We have not implemented TV option.\n");
      return MFISTA_E_TV_PRESET;
    }
  }
  return 0;
}

int mfista_imaging_core_nufft_batch(int nprob, const mfista_nufft_problem* pr, struct RESULT* res,
                                    mfista_opts* opts) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (nprob <= 0 || !pr || !res) return MFISTA_E_ARG;
  const int preset = opts ? opts->preset : MFISTA_PRESET_C24;
  int rc = 0;
  std::vector<int> tv;
  if ((rc = batch_check(nprob, pr, preset, &tv))) return rc;
  bool tuned = false;
  if (opts && opts->autotune) {
    int mmax = 0, reg = 0;
    for (int i = 0; i < nprob; i++) {
      mmax = std::max(mmax, pr[i].M);
      reg = std::max(reg, tune_reg(pr[i].lambda_tv, pr[i].lambda_tsv));
    }
    const TuneKey k = tune_key(pr[0].NX, pr[0].NY, preset, reg, mmax, nprob);
    TuneChoice ch;
    if (!tune_lookup(k, &ch)) {
      const int iters = opts->tune_iters > 0 ? opts->tune_iters : 30;
      auto run = [&](double* us) { return run_batch(C, nprob, pr, tv, opts, iters, false, nullptr, us); };
      if ((rc = tune_search(*C, reg, true, run, &ch, nullptr, toeplitz_batch_applies(nprob, pr, tv, opts))))
        return rc;
      tune_store(k, ch);
    }
    tune_apply(*C, ch);
    tuned = true;
  }
  rc = run_batch(C, nprob, pr, tv, opts, 0, true, res, nullptr);
  if (tuned) tune_restore(*C);
  return rc;
}

int mfista_autotune_nufft_batch(int nprob, const mfista_nufft_problem* pr, const mfista_opts* opts, int iters,
                                mfista_tune* best, mfista_tune* all, int max_all, int* n_all) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (nprob <= 0 || !pr || !best) return MFISTA_E_ARG;
  mfista_opts o;
  if (opts)
    o = *opts;
  else
    mfista_default_opts(&o);
  int rc = 0;
  std::vector<int> tv;
  if ((rc = batch_check(nprob, pr, o.preset, &tv))) return rc;
  int mmax = 0, reg = 0;
  for (int i = 0; i < nprob; i++) {
    mmax = std::max(mmax, pr[i].M);
    reg = std::max(reg, tune_reg(pr[i].lambda_tv, pr[i].lambda_tsv));
  }
  const int it = iters > 0 ? iters : 30;
  auto run = [&](double* us) { return run_batch(C, nprob, pr, tv, &o, it, false, nullptr, us); };
  TuneChoice ch;
  std::vector<TuneChoice> cand;
  if ((rc = tune_search(*C, reg, true, run, &ch, &cand, toeplitz_batch_applies(nprob, pr, tv, &o)))) return rc;
  const TuneKey k = tune_key(pr[0].NX, pr[0].NY, o.preset, reg, mmax, nprob);
  tune_store(k, ch);
  to_public(k, ch, best);
  for (int i = 0; all && i < (int)cand.size() && i < max_all; i++) to_public(k, cand[i], &all[i]);
  if (n_all) *n_all = (int)cand.size();
  return 0;
}

// Coupled movie, matlab/MFISTA_movie_eht.m:1-239: frames split by time_idx
// (1-based, :46-86), frame times t_list (:53-69), one persistent launch with
// one control step for all frames (solve_nufft_movie), then the outputs of
// :203-239 (model visibilities and the cost terms) from FP64 transforms.
int mfista_movie_nufft(const double* vis_r, const double* vis_i, const double* u_dx, const double* v_dy,
                       const double* noise_std, const int* time_idx, const double* time_val, int M,
                       const float* c_box, int NX, int NY, const double* xinit, double lambda_l1,
                       double lambda_tsv, double lambda_t, int nonneg_flag, double cinit, int maxiter,
                       double eps, double* xout, double* cost, double* terms, double* outvis_r,
                       double* outvis_i, int* iterations, mfista_opts* opts) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (M <= 0 || NX <= 0 || NY <= 0 || !vis_r || !vis_i || !u_dx || !v_dy || !noise_std || !time_idx ||
      !xout || maxiter <= 0 || lambda_t < 0)
    return MFISTA_E_ARG;
  const int preset = opts ? opts->preset : MFISTA_PRESET_C24;
  auto t0 = std::chrono::steady_clock::now();
  int nt = 0;
  for (int i = 0; i < M; i++) {
    if (time_idx[i] < 1) return MFISTA_E_ARG;
    nt = std::max(nt, time_idx[i]);
  }
  std::vector<std::vector<int>> fr(nt);
  for (int i = 0; i < M; i++) fr[time_idx[i] - 1].push_back(i);
  for (int f = 0; f < nt; f++)
    if (fr[f].empty()) return MFISTA_E_ARG;  // every frame needs data (its mean time, :62-64)
  std::vector<double> tl(nt);
  if (time_val) {
    for (int f = 0; f < nt; f++) {
      double sum = 0;
      for (int i : fr[f]) sum += time_val[i];
      tl[f] = sum / (double)fr[f].size();
    }
    const double span = *std::max_element(tl.begin(), tl.end()) - *std::min_element(tl.begin(), tl.end());
    if (!(span > 0)) return MFISTA_E_ARG;
    for (int f = 0; f < nt; f++) tl[f] = nt * tl[f] / span;
  } else {
    for (int f = 0; f < nt; f++) tl[f] = f + 1.0;
  }
  std::vector<double> w(nt > 1 ? nt - 1 : 0);
  for (int f = 0; f + 1 < nt; f++) {
    w[f] = 1.0 / (tl[f + 1] - tl[f]);
    if (!std::isfinite(w[f])) return MFISTA_E_ARG;
  }
  const size_t nn = (size_t)NX * NY, nb = nn * 8;
  std::vector<double> x0(nn, 0.0);
  for (size_t p = 0; p < nn; p++) {
    const double xv = xinit ? xinit[p] : 0.0;
    x0[p] = c_box ? xv * (double)c_box[p] : xv;  // x = (xinit .* c_box) for every frame, :49
  }
  if ((int)C->bset.size() < nt) C->bset.resize(nt);
  if ((int)C->bwork.size() < nt) C->bwork.resize(nt);
  int rc;
  std::vector<NufftState*> sets;
  std::vector<RunParams> P;
  std::vector<double> hb;
  for (int f = 0; f < nt; f++) {
    const int mf = (int)fr[f].size();
    BatchSet& S = C->bset[f];
    hb.resize((size_t)5 * mf);
    for (int k = 0; k < mf; k++) {
      const int i = fr[f][k];
      hb[k] = u_dx[i];
      hb[mf + k] = v_dy[i];
      hb[2 * mf + k] = vis_r[i];
      hb[3 * mf + k] = vis_i[i];
      hb[4 * mf + k] = noise_std[i];
    }
    const size_t Mb = (size_t)mf * 8;
    if ((rc = upload(*C, S.u, hb.data(), Mb)) || (rc = upload(*C, S.v, hb.data() + mf, Mb)) ||
        (rc = upload(*C, S.vr, hb.data() + 2 * mf, Mb)) || (rc = upload(*C, S.vi, hb.data() + 3 * mf, Mb)) ||
        (rc = upload(*C, S.sd, hb.data() + 4 * mf, Mb)))
      return rc;
    std::swap(C->nu, S.nu);
    rc = nufft_setup(*C, preset, (const double*)S.u.p, (const double*)S.v.p, (const double*)S.vr.p,
                     (const double*)S.vi.p, (const double*)S.sd.p, mf, NX, NY);
    std::swap(C->nu, S.nu);
    if (rc) {
      if (rc == MFISTA_E_FOV) printf("Error: the field of view is too small for the (u,v) coverage.\n");
      return rc;
    }
    PersistWork& W = C->bwork[f];
    if ((rc = upload(*C, W.x0, x0.data(), nb)) || (rc = upload(*C, W.xo, nullptr, nb))) return rc;
    if (c_box && (rc = upload(*C, W.box, c_box, nn * 4))) return rc;
    RunParams R;
    R.NX = NX;
    R.NY = NY;
    R.maxiter = maxiter;
    R.nonneg = (nonneg_flag == 1) ? 1 : 0;  // softth_nonneg only for nonneg_flag == 1, :141-146
    R.box_flag = c_box ? 1 : 0;
    R.tv_mode = 0;
    R.eps_abs = eps;  // the movie solver's rule is absolute and unscaled, :186
    R.rel_tol = opts ? opts->rel_tol : 0.0;
    R.lambda_l1 = lambda_l1;
    R.lambda_tsv = lambda_tsv;
    R.cinit = cinit;
    R.eta_first = kEta;
    R.single_trials = (opts && opts->single_trials) ? 1 : 0;
    R.deterministic = (opts && opts->deterministic) ? 1 : 0;
    R.xinit = (const double*)W.x0.p;
    R.xout = (double*)W.xo.p;
    R.box = c_box ? (const float*)W.box.p : nullptr;
    sets.push_back(&S.nu);
    P.push_back(R);
  }
  std::vector<RunStats> st;
  float kms = 0;
  if ((rc = solve_nufft_movie(*C, sets, P, lambda_t, w, st, &kms))) return rc;
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const int iters = st[0].ITER;
  if (iterations) *iterations = iters;
  if (cost && iters > 0) {
    MF_CHECK(hipMemcpyAsync(cost, C->bwork[0].cost.p, (size_t)std::min(iters, maxiter) * 8, hipMemcpyDeviceToHost,
                            C->stream));
    MF_CHECK(hipStreamSynchronize(C->stream));
  }
  // outputs (:203-239): images, FP64 chi^2 / L1 / TSV per frame, d_X, model visibilities
  double chisq = 0, l1 = 0, tsvc = 0;
  std::vector<double> model;
  for (int f = 0; f < nt; f++) {
    PersistWork& W = C->bwork[f];
    MF_CHECK(hipMemcpyAsync(xout + (size_t)f * nn, W.xo.p, nb, hipMemcpyDeviceToHost, C->stream));
    MF_CHECK(hipStreamSynchronize(C->stream));
    NufftState& N = C->bset[f].nu;
    std::swap(C->nu, N);
    double chi2 = 0, stats[4] = {0, 0, 0, 0};
    rc = nufft_chi2_64(*C, (const double*)W.xo.p, &chi2);
    if (!rc) rc = image_stats64(*C, (const double*)W.xo.p, NX, NY, stats);
    if (!rc && outvis_r && outvis_i) {
      const int mf = (int)fr[f].size();
      if (C->h_model.ensure((size_t)mf * sizeof(double2))) {
        rc = MFISTA_E_HIP;
      } else {
        nufft_fwd<double>(*C, (const double*)W.xo.p, nullptr, nullptr, (double2*)C->h_model.p, nullptr);
        model.resize((size_t)2 * mf);
        // the copy is ordered after the transform on the library's stream (a non-blocking stream)
        if (hipMemcpyAsync(model.data(), C->h_model.p, (size_t)mf * 16, hipMemcpyDeviceToHost, C->stream) !=
                hipSuccess ||
            hipStreamSynchronize(C->stream) != hipSuccess)
          rc = MFISTA_E_HIP;
        for (int k = 0; k < mf && !rc; k++) {
          outvis_r[fr[f][k]] = model[2 * k];
          outvis_i[fr[f][k]] = model[2 * k + 1];
        }
      }
    }
    std::swap(C->nu, N);
    if (rc) return rc;
    chisq += chi2;
    l1 += stats[0];
    tsvc += stats[1];
  }
  double dx = 0;
  for (int f = 0; f + 1 < nt; f++) {
    const double* a = xout + (size_t)f * nn;
    const double* b = a + nn;
    double s1 = 0;
    for (size_t p = 0; p < nn; p++) s1 += (a[p] - b[p]) * (a[p] - b[p]);
    dx += w[f] * s1;
  }
  if (terms) {
    terms[0] = chisq;
    terms[1] = l1;
    terms[2] = tsvc;
    terms[3] = dx;
  }
  if (opts) {
    int tt = 0, mt = 0;
    for (const RunStats& r : st) mt = std::max(mt, r.max_trials);
    tt = st[0].total_trials;
    opts->engine_used = MFISTA_ENGINE_PERSISTENT;
    opts->solve_ms = secs * 1e3;
    opts->kernel_ms = kms;
    opts->total_trials = tt;
    opts->max_trials = mt;
    opts->trial_rounds = st[0].rounds;
  }
  return st[0].status ? MFISTA_E_HIP : 0;
}

int mfista_imaging_core_nufft_ex(const double* u_dx, const double* v_dy, const double* vis_r,
                                 const double* vis_i, const double* vis_std, int M, int NX,
                                 int NY, int maxiter, double eps, double lambda_l1,
                                 double lambda_tv, double lambda_tsv, double cinit,
                                 const double* xinit, double* xout, int nonneg_flag,
                                 int box_flag, const float* cl_box, struct RESULT* res,
                                 mfista_opts* opts) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (M <= 0 || NX <= 0 || NY <= 0 || !u_dx || !v_dy || !vis_r || !vis_i || !vis_std || !xinit ||
      !xout || !res)
    return MFISTA_E_ARG;
  const size_t Mb = (size_t)M * 8, nb = (size_t)NX * NY * 8;
  int rc;
  if ((rc = upload(*C, C->h_u, u_dx, Mb)) || (rc = upload(*C, C->h_v, v_dy, Mb)) ||
      (rc = upload(*C, C->h_vr, vis_r, Mb)) || (rc = upload(*C, C->h_vi, vis_i, Mb)) ||
      (rc = upload(*C, C->h_sd, vis_std, Mb)) || (rc = upload(*C, C->h_x0, xinit, nb)) ||
      (rc = upload(*C, C->h_x1, nullptr, nb)))
    return rc;
  if (box_flag) {
    if (!cl_box) return MFISTA_E_ARG;
    if ((rc = upload(*C, C->h_box, cl_box, (size_t)NX * NY * 4))) return rc;
  }
  SolveArgs A;
  A.u = (const double*)C->h_u.p;
  A.v = (const double*)C->h_v.p;
  A.vr = (const double*)C->h_vr.p;
  A.vi = (const double*)C->h_vi.p;
  A.sd = (const double*)C->h_sd.p;
  A.M = M;
  A.NX = NX;
  A.NY = NY;
  A.maxiter = maxiter;
  A.eps = eps;
  A.lambda_l1 = lambda_l1;
  A.lambda_tv = lambda_tv;
  A.lambda_tsv = lambda_tsv;
  A.cinit = cinit;
  A.xinit = (const double*)C->h_x0.p;
  A.xout = (double*)C->h_x1.p;
  A.nonneg = nonneg_flag;
  A.box_flag = box_flag;
  A.box = box_flag ? (const float*)C->h_box.p : nullptr;
  RESULT tmp = *res;
  rc = solve_nufft(*C, A, &tmp, opts);
  if (rc) return rc;
  if (hipMemcpy(xout, C->h_x1.p, nb, hipMemcpyDeviceToHost) != hipSuccess) return MFISTA_E_HIP;
  *res = tmp;
  return 0;
}

int mfista_autotune_nufft(const double* u_dx, const double* v_dy, const double* vis_r,
                          const double* vis_i, const double* vis_std, int M, int NX, int NY,
                          double lambda_l1, double lambda_tv, double lambda_tsv, double cinit,
                          const double* xinit, int nonneg_flag, const mfista_opts* opts, int iters,
                          mfista_tune* best, mfista_tune* all, int max_all, int* n_all) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (M <= 0 || NX <= 0 || NY <= 0 || !u_dx || !v_dy || !vis_r || !vis_i || !vis_std || !best)
    return MFISTA_E_ARG;
  const int preset = opts ? opts->preset : MFISTA_PRESET_C24;
  if (lambda_tv > 0 && lambda_tsv > 0) return MFISTA_E_BOTH_REG;
  if (lambda_tv > 0 && preset == MFISTA_PRESET_CXX12) return MFISTA_E_TV_PRESET;
  const size_t Mb = (size_t)M * 8, nb = (size_t)NX * NY * 8;
  int rc;
  if ((rc = upload(*C, C->h_u, u_dx, Mb)) || (rc = upload(*C, C->h_v, v_dy, Mb)) ||
      (rc = upload(*C, C->h_vr, vis_r, Mb)) || (rc = upload(*C, C->h_vi, vis_i, Mb)) ||
      (rc = upload(*C, C->h_sd, vis_std, Mb)) || (rc = upload(*C, C->h_x0, xinit, nb)) ||
      (rc = upload(*C, C->h_x1, nullptr, nb)))
    return rc;
  if (!xinit) MF_CHECK(hipMemsetAsync(C->h_x0.p, 0, nb, C->stream));
  SolveArgs A;
  A.u = (const double*)C->h_u.p;
  A.v = (const double*)C->h_v.p;
  A.vr = (const double*)C->h_vr.p;
  A.vi = (const double*)C->h_vi.p;
  A.sd = (const double*)C->h_sd.p;
  A.M = M;
  A.NX = NX;
  A.NY = NY;
  A.lambda_l1 = lambda_l1;
  A.lambda_tv = lambda_tv;
  A.lambda_tsv = lambda_tsv;
  A.cinit = cinit;
  A.xinit = (const double*)C->h_x0.p;
  A.xout = (double*)C->h_x1.p;
  A.nonneg = nonneg_flag;
  TuneChoice ch;
  std::vector<TuneChoice> cand;
  if ((rc = tune_single(*C, A, opts, iters, &ch, &cand))) return rc;
  const TuneKey k = tune_key(NX, NY, preset, tune_reg(lambda_tv, lambda_tsv), M, 1);
  tune_store(k, ch);
  to_public(k, ch, best);
  for (int i = 0; all && i < (int)cand.size() && i < max_all; i++) to_public(k, cand[i], &all[i]);
  if (n_all) *n_all = (int)cand.size();
  return 0;
}

void mfista_imaging_core_nufft(double* u_dx, double* v_dy, double* vis_r, double* vis_i,
                               double* vis_std, int M, int NX, int NY, int maxiter, double eps,
                               double lambda_l1, double lambda_tv, double lambda_tsv,
                               double cinit, double* xinit, double* xout, int nonneg_flag,
                               int box_flag, float* cl_box, struct RESULT* mfista_result) {
  mfista_opts o;
  mfista_default_opts(&o);
  int rc = mfista_imaging_core_nufft_ex(u_dx, v_dy, vis_r, vis_i, vis_std, M, NX, NY, maxiter,
                                        eps, lambda_l1, lambda_tv, lambda_tsv, cinit, xinit, xout,
                                        nonneg_flag, box_flag, cl_box, mfista_result, &o);
  if (rc && rc != MFISTA_E_BOTH_REG)
    fprintf(stderr, "[mfista_amd] mfista_imaging_core_nufft failed with status %d\n", rc);
}

static int nufft_operator(int preset, int fp64, const double* u, const double* v, int M, int NX,
                          int NY, const double* x_or_null, const double* yr, const double* yi,
                          double* out_yr, double* out_yi, double* out_x) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (M <= 0 || NX <= 0 || NY <= 0 || !u || !v) return MFISTA_E_ARG;
  const size_t Mb = (size_t)M * 8, nb = (size_t)NX * NY * 8;
  int rc;
  if ((rc = upload(*C, C->h_u, u, Mb)) || (rc = upload(*C, C->h_v, v, Mb))) return rc;
  if ((rc = nufft_setup(*C, preset, (const double*)C->h_u.p, (const double*)C->h_v.p, nullptr,
                        nullptr, nullptr, M, NX, NY)))
    return rc;
  NufftState& N = C->nu;
  hipStream_t st = C->stream;
  if (fp64 && nufft_ensure64(*C)) return MFISTA_E_HIP;
  if (x_or_null) {  // forward
    if ((rc = upload(*C, C->h_x0, x_or_null, nb))) return rc;
    if (C->h_model.ensure((size_t)M * sizeof(double2)) || C->h_vr.ensure(Mb) ||
        C->h_vi.ensure(Mb) || C->x.ensure(nb))
      return MFISTA_E_HIP;
    if (fp64) {
      nufft_fwd<double>(*C, (const double*)C->h_x0.p, nullptr, nullptr, (double2*)C->h_model.p,
                        nullptr);
      hipLaunchKernelGGL(k_split_model<double>, dim3((M + 255) / 256), dim3(256), 0, st,
                         (const double2*)C->h_model.p, (double*)C->h_vr.p, (double*)C->h_vi.p, M);
    } else {
      launch_cast<double, float>((const double*)C->h_x0.p, (float*)C->x.p, NX * NY, st);
      nufft_fwd<float>(*C, (const float*)C->x.p, nullptr, nullptr, (float2*)C->h_model.p,
                       nullptr);
      hipLaunchKernelGGL(k_split_model<float>, dim3((M + 255) / 256), dim3(256), 0, st,
                         (const float2*)C->h_model.p, (double*)C->h_vr.p, (double*)C->h_vi.p, M);
    }
    MF_CHECK(hipMemcpyAsync(out_yr, C->h_vr.p, Mb, hipMemcpyDeviceToHost, st));
    MF_CHECK(hipMemcpyAsync(out_yi, C->h_vi.p, Mb, hipMemcpyDeviceToHost, st));
  } else {  // adjoint
    if ((rc = upload(*C, C->h_vr, yr, Mb)) || (rc = upload(*C, C->h_vi, yi, Mb))) return rc;
    if (C->h_x1.ensure(nb) || C->x.ensure(nb)) return MFISTA_E_HIP;
    if (fp64) {
      hipLaunchKernelGGL(k_gather_vis<double>, dim3((M + 255) / 256), dim3(256), 0, st,
                         (const double*)C->h_vr.p, (const double*)C->h_vi.p,
                         (const int*)N.perm.p, (const int*)N.yneg.p, M, (double2*)N.rw64.p);
      nufft_adj<double>(*C, (const double2*)N.rw64.p, (double*)C->h_x1.p, nullptr);
    } else {
      hipLaunchKernelGGL(k_gather_vis<float>, dim3((M + 255) / 256), dim3(256), 0, st,
                         (const double*)C->h_vr.p, (const double*)C->h_vi.p,
                         (const int*)N.perm.p, (const int*)N.yneg.p, M, (float2*)N.rw.p);
      nufft_adj<float>(*C, (const float2*)N.rw.p, (float*)C->x.p, nullptr);
      launch_cast<float, double>((const float*)C->x.p, (double*)C->h_x1.p, NX * NY, st);
    }
    MF_CHECK(hipMemcpyAsync(out_x, C->h_x1.p, nb, hipMemcpyDeviceToHost, st));
  }
  MF_CHECK(hipStreamSynchronize(st));
  MF_CHECK(hipGetLastError());
  return 0;
}

int mfista_nufft_forward(int preset, int fp64, const double* u, const double* v, int M, int NX,
                         int NY, const double* x, double* y_r, double* y_i) {
  if (!x || !y_r || !y_i) return MFISTA_E_ARG;
  return nufft_operator(preset, fp64, u, v, M, NX, NY, x, nullptr, nullptr, y_r, y_i, nullptr);
}

int mfista_nufft_adjoint(int preset, int fp64, const double* u, const double* v, int M, int NX,
                         int NY, const double* y_r, const double* y_i, double* x) {
  if (!y_r || !y_i || !x) return MFISTA_E_ARG;
  return nufft_operator(preset, fp64, u, v, M, NX, NY, nullptr, y_r, y_i, nullptr, nullptr, x);
}

void x2y_nufft(double* u_dx, double* v_dy, int M, int Nx, int Ny, double* xin, double* y_r,
               double* y_i) {
  int rc = mfista_nufft_forward(MFISTA_PRESET_CXX12, 1, u_dx, v_dy, M, Nx, Ny, xin, y_r, y_i);
  if (rc) fprintf(stderr, "[mfista_amd] x2y_nufft failed with status %d\n", rc);
}

void y2x_nufft(double* u_dx, double* v_dy, int M, int Nx, int Ny, double* yin_r, double* yin_i,
               double* xout) {
  int rc = mfista_nufft_adjoint(MFISTA_PRESET_CXX12, 1, u_dx, v_dy, M, Nx, Ny, yin_r, yin_i, xout);
  if (rc) fprintf(stderr, "[mfista_amd] y2x_nufft failed with status %d\n", rc);
}


// per-kernel timing of the last NUFFT solve's kernels, one event pair per
// launch on the library stream (the tracing hook bench.py reads)
int mfista_profile_stages(int reps, double* ms, int nstage) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C || !C->nu.ready || reps <= 0 || !ms) return MFISTA_E_ARG;
  NufftState& N = C->nu;
  hipStream_t st = C->stream;
  const int NX = N.NX, NY = N.NY, n = NX * NY;
  const GridGeom& g = N.geom;
  // the stream-engine buffers these kernels use, allocated here: the last
  // solve may have run on the persistent engine, which has its own
  {
    const size_t rb = (size_t)n * 4;
    const size_t np = (size_t)degrid_blocks(N.M, N.MSP) + (size_t)(kProxW + 1) * elem_blocks(n);
    if (C->z.ensure(rb) || C->xnew.ensure(rb) || C->g.ensure(rb) || C->dstep.ensure(rb) ||
        C->part.ensure(np * 8) || C->state.ensure(sizeof(DevState)))
      return MFISTA_E_HIP;
    MF_CHECK(hipMemsetAsync(C->z.p, 0, rb, st));
    MF_CHECK(hipMemsetAsync(C->g.p, 0, rb, st));
    DevState h;
    memset(&h, 0, sizeof(h));
    h.c = 1.0;
    h.lambda_l1 = 1.0;
    h.nonneg = 1;
    MF_CHECK(hipMemcpy(C->state.p, &h, sizeof(h), hipMemcpyHostToDevice));
  }
  float* z = (float*)C->z.p;
  float* xnew = (float*)C->xnew.p;
  float* gr = (float*)C->g.p;
  double* part = (double*)C->part.p;
  const DevState* S = (const DevState*)C->state.p;
  std::vector<hipEvent_t> e0(reps), e1(reps);
  for (int r = 0; r < reps; r++) {
    MF_CHECK(hipEventCreate(&e0[r]));
    MF_CHECK(hipEventCreate(&e1[r]));
  }
  auto run = [&](int stage) {
    switch (stage) {
      case 0:
        launch_rowfft_fwd<float>(z, (const float*)N.e4xf.p, (const float*)N.e4yf.p, (float2*)N.T.p, NX,
                                 NY, g.Ly, NY / 2, N.Py, (const float2*)N.twyf.p, st, nullptr);
        break;
      case 1:
        launch_colfft_fwd<float>((const float2*)N.T.p, (float2*)N.Gf.p, NX, NY + 1, g.Lx, NX / 2, 0,
                                 N.MSP, N.Px, (const float2*)N.twxf.p, st, nullptr);
        break;
      case 2:
        launch_degrid<float>(g, N.pf(), (const float2*)N.Gf.p, nullptr, (float2*)N.rw.p, nullptr, part,
                             nullptr, 0, degrid_blocks(N.M, N.MSP), nullptr, nullptr, nullptr, st,
                             nullptr);
        break;
      case 3:
        launch_grid<float>(g, N.pf(), (const float2*)N.rw.p, (const int4*)N.chunks.p, N.nchunks,
                           (const unsigned*)N.ent.p, (float2*)N.Hg.p, st, nullptr);
        break;
      case 4:
        launch_colfft_inv<float>((float2*)N.Hg.p, (float2*)N.U.p, NX, NY + 1, g.Lx, NX / 2, 0,
                                 N.Px, (const float2*)N.twxf.p, st, nullptr);
        break;
      case 5:
        launch_rowfft_c2r<float>((const float2*)N.U.p, (const float*)N.e4xf.p, (const float*)N.e4yf.p,
                                 1.0f / (8.0f * NX * NY), gr, NX, NY, g.Ly, NY / 2, N.Py,
                                 (const float2*)N.twyf.p, st, nullptr);
        break;
      case 6:
        launch_tsv_grad<float>(z, gr, NX, NY, 0.0, part, st, nullptr);
        break;
      case 7:
        launch_prox<float>(z, gr, xnew, (float*)C->dstep.p, nullptr, NX, NY, S, part, st, nullptr);
        break;
    }
  };
  const int ns = nstage < 8 ? nstage : 8;
  for (int stage = 0; stage < ns; stage++) {
    run(stage);
    run(stage);
    for (int r = 0; r < reps; r++) {
      MF_CHECK(hipEventRecord(e0[r], st));
      run(stage);
      MF_CHECK(hipEventRecord(e1[r], st));
    }
    MF_CHECK(hipStreamSynchronize(st));
    double tot = 0;
    for (int r = 0; r < reps; r++) {
      float t = 0;
      MF_CHECK(hipEventElapsedTime(&t, e0[r], e1[r]));
      tot += t;
    }
    ms[stage] = tot / reps;
  }
  for (int r = 0; r < reps; r++) {
    (void)hipEventDestroy(e0[r]);
    (void)hipEventDestroy(e1[r]);
  }
  MF_CHECK(hipGetLastError());
  return 0;
}

long mfista_last_steps(void) {
  Context* C = get_context();
  return C ? C->last_steps : -1;
}

}  // extern "C"
