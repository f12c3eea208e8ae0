// wide.hip -- the batched-kernel ("wide") engine for many NUFFT solves at once.
//
// The loop is the stream engine's (driver.hip mfista_run, restating
// c/mfista_nufft_lib.c:557-787): every step is one backtracking trial,
// preceded by the iteration prologue (FP32 adjoint of the residual at z, TSV
// gradient) when the previous trial was accepted, and a one-workgroup FP64
// control step per problem sets the gates of the next step.  As in the
// persistent engine, the residual at z needs no forward transform: the
// operator is linear, so with FP64 weighted models m_x = W A x and
// m_xn = W A xnew, W A z = a m_x + b m_xn for z = a x + b xnew; each trial
// computes W A d for its step d = xnew - z with an FP64 forward transform and
// m_xn = m_z + W A d.  What changes is the launch shape: every kernel of a step covers all
// problems of the batch at once (blockIdx.y = problem, blockIdx.x = the
// stage's task within it), so a stage of 32 problems at 256^2 fills the GPU
// with thousands of workgroups instead of a hundred, and the kernels keep
// their own small register footprints (several workgroups per CU), where the
// persistent batch shares one 255-VGPR kernel across all stages.  A problem
// whose gates are closed (converged, or accepting in the middle of a trial
// sequence) returns at once from the kernels that do not apply to it.
//
// The step's launch sequence is captured once per batch in a hipGraph and
// replayed; the host reads one counter of finished problems every few steps.
// Everything a kernel needs is in device memory (per-problem descriptors and
// states), so replays need no host work.  No memset is captured: the adjoint
// grids are cleared by the column pass that reads them.
//
// Toeplitz steps (toeplitz_body.h): for L1 / L1+TSV problems on square images
// the data term's normal operator is applied as an FP64 FFT convolution, so
// the images T x, T xnew and T z replace the visibility-domain models, the
// gradient at z is b - T z (one elementwise kernel instead of the gridding
// adjoint) and a trial's forward transform is the row FFT of its step, the
// column product with Kh and the c2r (no degridding).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "driver.h"
#include "mfista_dev.h"
#include "solver.h"
#include "stages.h"
#include "toeplitz_body.h"

namespace mf {

// One problem of a wide batch (device copy; the kernels index it by blockIdx.y).
struct WProb {
  GridGeom g;  // of the problem's visibility set (M, tiles)
  VisParams<double> pd;
  VisParams<float> pf;
  const int4* chunks;  // gridding chunk list of the set
  const unsigned* ent;
  int nchunks, n_degrid, M;
  PartialLayout L;
  double *x, *z, *xnew, *gr, *d;  // FP64 iterates, gradient and step
  double2 *T64, *Gf64;            // FP64 forward transform (of the step or the start)
  double2 *mx, *mxn, *mz;         // FP64 weighted models W A x, W A xnew, W A z (sorted order)
  float2* rz32;                   // FP32 weighted residual W V - W A z (gridding input)
  float2 *Hg, *U;                 // FP32 adjoint grid (kept zero between adjoints) and column pass
  double *part, *stat, *cost;
  DevState* S;
  const float* box;
  double lambda_tsv;
  // Toeplitz steps: the set's kernel spectrum and b, and T x, T xnew, T z
  const double *kh, *bimg;
  double *tx, *txn, *tz;
};

// What the problems of a batch share: image size, preset, FFT plans, twiddle
// and E4 tables (they depend only on the size and the preset).
struct WShared {
  FFTPlan Px, Py;
  const double2 *twx64, *twy64;
  const float2 *twx32, *twy32;
  const double *e4x64, *e4y64;
  const float *e4x32, *e4y32;
  int NX, NY, Lx, Ly, MSP, TX, TY;
  int nseq_r64, nseq_c64, nseq_ci32, nseq_c2r32;
  float adj_scale;  // 1 / (8 NX NY)
  int* ndone;       // problems finished
  FFTPlan Pxh;      // Toeplitz steps: the NX-point plan and twiddles of the pruned transforms
  const double2* twxh64;
};

// which transform a forward kernel computes: of the start (no gate) or of a
// trial's step (gate_trial)
enum WFwd { kWInit = 0, kWTrial = 2 };

__device__ inline bool w_closed(const DevState* S, int which) {
  return which == kWTrial && S->gate_trial != 0;
}

__global__ void __launch_bounds__(MF_BLOCK) kw_rowfft64(const WProb* __restrict__ W, WShared sh, int which) {
  const WProb& p = W[blockIdx.y];
  if (w_closed(p.S, which)) return;
  extern __shared__ double4 smem_d4[];
  const double* x = (which == kWInit) ? p.x : p.d;
  const int NY = sh.NY;
  auto load = [&](int i, int j) -> double { return x[(size_t)i * NY + j] * sh.e4x64[i] * sh.e4y64[j]; };
  st_rowfft_fwd<double>(blockIdx.x, sh.nseq_r64, sh.NX, NY, sh.Ly, NY / 2, sh.Py, sh.twy64, p.T64,
                        reinterpret_cast<double2*>(smem_d4), load);
}

__global__ void __launch_bounds__(MF_BLOCK) kw_colfft64(const WProb* __restrict__ W, WShared sh, int which) {
  const WProb& p = W[blockIdx.y];
  if (w_closed(p.S, which)) return;
  extern __shared__ double4 smem_d4[];
  st_colfft_fwd<double>(blockIdx.x, sh.nseq_c64, p.T64, p.Gf64, sh.NX, sh.NY + 1, sh.Lx, sh.NX / 2, 0, sh.MSP,
                        sh.Px, sh.twx64, reinterpret_cast<double2*>(smem_d4));
}

// FP64 degridding into the weighted models.
//   init:  t = W A x0; m_x = t; chi^2 partials of the start (L.chi_x)
//   trial: t = W A d; m_xn = m_z + t; partials of |W V - m_z - t|^2 (chi^2
//          at xnew, L.chi_x) and of |t|^2 (L.lip)
template <int MSPC>
__global__ void __launch_bounds__(MF_BLOCK) kw_degrid64(const WProb* __restrict__ W, WShared sh, int which) {
  const WProb& p = W[blockIdx.y];
  if ((int)blockIdx.x >= p.n_degrid || w_closed(p.S, which)) return;
  __shared__ double red[32];
  double v[2] = {0.0, 0.0};
  int s;
  double2 F;
  if (degrid_core<double, MSPC>(blockIdx.x, p.g, p.pd, p.Gf64, s, F)) {
    const double w = p.pd.w[s];
    const double2 V = p.pd.vis[s];
    const double tx = F.x * w, ty = F.y * w;
    if (which == kWTrial) {
      const double2 mz = p.mz[s];
      const double rx = V.x * w - mz.x - tx, ry = V.y * w - mz.y - ty;
      v[0] = rx * rx + ry * ry;
      v[1] = tx * tx + ty * ty;
      p.mxn[s] = make_double2(mz.x + tx, mz.y + ty);
    } else {
      const double rx = V.x * w - tx, ry = V.y * w - ty;
      v[0] = rx * rx + ry * ry;
      p.mx[s] = make_double2(tx, ty);
    }
  }
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    p.part[p.L.chi_x + blockIdx.x] = v[0];
    if (which == kWTrial) p.part[p.L.lip + blockIdx.x] = v[1];
  }
}

// Momentum in the visibility domain, after the control step of an iteration
// that ends (gate_momentum open), or at the start (init: z = x):
// m_z = a m_x + b m_xn, m_x <- m_xn on acceptance; the FP32 residual
// W V - m_z for the next adjoint and the partials of its chi^2 (L.chi_z).
__device__ inline void vis_momentum(const WProb& p, int blk, int init) {
  const DevState* S = p.S;
  const int nvb = (p.M + MF_BLOCK - 1) / MF_BLOCK;
  if (blk >= nvb || (!init && S->gate_momentum)) return;
  __shared__ double red[16];
  const int s = blk * MF_BLOCK + threadIdx.x;
  double v[1] = {0.0};
  if (s < p.M) {
    double2 mz;
    if (init) {
      mz = p.mx[s];
    } else {
      const double a = S->mom_a, b = S->mom_b;
      const double2 mx = p.mx[s], mn = p.mxn[s];
      mz = make_double2(a * mx.x + b * mn.x, a * mx.y + b * mn.y);
      if (S->accept_x) p.mx[s] = mn;
    }
    p.mz[s] = mz;
    const double w = p.pd.w[s];
    const double2 V = p.pd.vis[s];
    const double rx = V.x * w - mz.x, ry = V.y * w - mz.y;
    v[0] = rx * rx + ry * ry;
    p.rz32[s] = make_float2((float)rx, (float)ry);
  }
  block_sum<1>(v, red);
  if (threadIdx.x == 0) p.part[p.L.chi_z + blk] = v[0];
}

__global__ void __launch_bounds__(MF_BLOCK) kw_vis_momentum(const WProb* __restrict__ W, int init) {
  vis_momentum(W[blockIdx.y], blockIdx.x, init);
}

// Gridding with separable tap weights.  The Gaussian weight of tap (l, l') is
// exp(l a_x - l^2 t3x) exp(l' a_y - l'^2 t3y), so for one entry and one tile
// every column c has one x weight and every row r one y weight (zero outside
// the 2 MSP footprint).  Per batch of kSepEB entries, the first half of the
// workgroup tabulates the column weights and the deposit of entry q, the
// second half its row weights (TX + TY exponentials per entry instead of
// one per covered cell); each thread then adds wx[c] wy[r] dep to its cells.
// Tile bookkeeping and flushes are grid_chunks'.
constexpr int kSepChunks = 4;                         // chunks per batch
constexpr int kSepEB = kSepChunks * kGridChunk;       // entries per batch (128)

__host__ __device__ inline size_t grid_sep_lds_bytes(int TX, int TY) {
  return (size_t)kSepEB * (TX + TY + 2) * sizeof(float) + kSepChunks * sizeof(int4);
}

template <int CPT, typename Resid>
__device__ inline void grid_sep(const GridGeom& g, const VisParams<float>& vp, Resid resid,
                                const int4* __restrict__ chunks, int cb, int ce, const unsigned* __restrict__ ent,
                                float2* __restrict__ Hg, char* lds) {
  const int TX = g.TX, TY = g.TY, SW = TX + TY + 2;
  float* tab = reinterpret_cast<float*>(lds);
  int4* cmeta = reinterpret_cast<int4*>(tab + (size_t)kSepEB * SW);
  const int NX = g.NX, MSP = g.MSP;
  const double t3x = g.t3x, t3y = g.t3y;
  constexpr double kL2E = 1.4426950408889634074;  // log2(e)
  int rr[CPT], cc[CPT];
#pragma unroll
  for (int c = 0; c < CPT; c++) {
    const int cell = threadIdx.x + c * MF_BLOCK;
    rr[c] = cell / TX;
    cc[c] = cell - rr[c] * TX;
  }
  float2 acc[CPT];
#pragma unroll
  for (int c = 0; c < CPT; c++) acc[c] = make_float2(0.f, 0.f);
  int cur = -1, cur_first = 0, rb = 0, cbx = 0;
  auto flush = [&](bool owned) {
#pragma unroll
    for (int c = 0; c < CPT; c++) {
      const int row = rb + rr[c], col = cbx + cc[c];
      const float2 v = acc[c];
      if (rr[c] < TY && row <= g.NY && col < g.Lx && (v.x != 0.f || v.y != 0.f)) {
        float2* dst = Hg + (size_t)row * g.Lx + col;
        if (owned) {
          *dst = v;
        } else {
          unsafeAtomicAdd(&dst->x, v.x);
          unsafeAtomicAdd(&dst->y, v.y);
        }
      }
      acc[c] = make_float2(0.f, 0.f);
    }
  };
  for (int c0 = cb; c0 < ce; c0 += kSepChunks) {
    const int nc = min(kSepChunks, ce - c0);
    __syncthreads();  // the previous batch's tables are consumed
    {
      const int q = threadIdx.x & (kSepEB - 1), half = threadIdx.x / kSepEB, j = q / kGridChunk;
      if (j < nc) {
        const int4 ch = chunks[c0 + j];
        if (half == 0 && (q % kGridChunk) == 0) cmeta[j] = ch;
        const int e = ch.y + (q % kGridChunk);
        if (e < ch.z) {
          const unsigned en = ent[e];
          const int sv = (int)(en & 0x7fffffffu);
          const bool mir = (en >> 31) != 0;
          const int2 m = GP(vp.mxy)[sv];
          const float2 a = GP(vp.a)[sv];
          const int tr = ch.x / g.ntx, tc = ch.x - tr * g.ntx;
          const int trb = tr * TY, tcb = tc * TX;
          // tap offsets of cell (row, col): direct l' = row - my, l = col - mx - NX;
          // mirror l' = -row - my, l = NX - mx - col
          const int sg = mir ? -1 : 1;
          float* rec = tab + (size_t)q * SW;
          if (half == 0) {
            const int ox = mir ? NX - m.x - tcb : tcb - m.x - NX;
            const double ax = a.x;
            for (int c = 0; c < TX; c++) {
              const int l = ox + sg * c;
              rec[c] = (l > -MSP && l <= MSP) ? __builtin_amdgcn_exp2f((float)(kL2E * (l * ax - (double)l * l * t3x)))
                                              : 0.f;
            }
            const float ew = GP(vp.e1)[sv] * GP(vp.w)[sv];
            const float2 r = resid(sv);
            rec[TX + TY] = ew * r.x;
            rec[TX + TY + 1] = mir ? ew * r.y : -ew * r.y;
          } else {
            const int oy = mir ? -(trb + m.y) : trb - m.y;
            const double ay = a.y;
            for (int r = 0; r < TY; r++) {
              const int l = oy + sg * r;
              rec[TX + r] = (l > -MSP && l <= MSP)
                                ? __builtin_amdgcn_exp2f((float)(kL2E * (l * ay - (double)l * l * t3y)))
                                : 0.f;
            }
          }
        }
      }
    }
    __syncthreads();
    for (int j = 0; j < nc; j++) {
      const int4 ch = cmeta[j];
      if (ch.x != cur) {
        if (cur >= 0) flush(false);
        cur = ch.x;
        cur_first = ch.w & 1;
        const int tr = cur / g.ntx, tc = cur - tr * g.ntx;
        rb = tr * TY;
        cbx = tc * TX;
      }
      const int q0 = j * kGridChunk, q1 = q0 + min(kGridChunk, ch.z - ch.y);
#pragma unroll 4
      for (int q = q0; q < q1; q++) {
        const float* rec = tab + (size_t)q * SW;
        const float dx = rec[TX + TY], dy = rec[TX + TY + 1];
#pragma unroll
        for (int c = 0; c < CPT; c++) {
          const float w = rec[cc[c]] * rec[TX + rr[c]];
          acc[c].x = fmaf(w, dx, acc[c].x);
          acc[c].y = fmaf(w, dy, acc[c].y);
        }
      }
      if (ch.w & 2) {  // last chunk of the tile
        flush(cur_first != 0);
        cur = -1;
      }
    }
  }
  if (cur >= 0) flush(false);
  __syncthreads();
}

// Gridding of the FP32 residual: workgroup x takes chunks [kGB x, kGB x + kGB)
// of the problem's tile-ordered chunk list (each thread owns CPT cells of the
// tile and sums in registers; a tile whose chunks all fall in this range is
// stored, others are flushed with atomics).
template <int CPT>
__global__ void __launch_bounds__(MF_BLOCK) kw_grid32(const WProb* __restrict__ W, WShared sh) {
  const WProb& p = W[blockIdx.y];
  const int cb = blockIdx.x * kGB;
  if (cb >= p.nchunks || p.S->gate_prologue) return;
  extern __shared__ double4 smem_d4[];
  const float2* rz = p.rz32;
  grid_sep<CPT>(p.g, p.pf, [&](int s) { return rz[s]; }, p.chunks, cb, min(cb + kGB, p.nchunks), p.ent, p.Hg,
                reinterpret_cast<char*>(smem_d4));
}

__global__ void __launch_bounds__(MF_BLOCK) kw_colinv32(const WProb* __restrict__ W, WShared sh) {
  const WProb& p = W[blockIdx.y];
  if (p.S->gate_prologue) return;
  extern __shared__ double4 smem_d4[];
  st_colfft_inv<float>(blockIdx.x, sh.nseq_ci32, p.Hg, p.U, sh.NX, sh.NY + 1, sh.Lx, sh.NX / 2, 0, sh.Px,
                       sh.twx32, reinterpret_cast<float2*>(smem_d4));
}

// c2r along ky, crop, E4 / (8 NX NY) in FP32 (as the stream engine's FP32
// adjoint), stored as the FP64 gradient
__global__ void __launch_bounds__(MF_BLOCK) kw_c2r32(const WProb* __restrict__ W, WShared sh) {
  const WProb& p = W[blockIdx.y];
  if (p.S->gate_prologue) return;
  extern __shared__ double4 smem_d4[];
  const int NY = sh.NY;
  double* gr = p.gr;
  auto store = [&](int i, int j, float v) {
    v *= sh.adj_scale;
    v *= sh.e4x32[i];
    v *= sh.e4y32[j];
    gr[(size_t)i * NY + j] = (double)v;
  };
  st_rowfft_c2r<float>(blockIdx.x, sh.nseq_c2r32, p.U, sh.NX, NY, sh.Ly, NY / 2, sh.Py, sh.twy32,
                       reinterpret_cast<float2*>(smem_d4), store);
}

__global__ void __launch_bounds__(MF_BLOCK) kw_tsv_grad(const WProb* __restrict__ W, WShared sh) {
  const WProb& p = W[blockIdx.y];
  if (p.S->gate_prologue || !(p.lambda_tsv > 0.0)) return;
  __shared__ double red[16];
  dev_tsv_grad<double>(blockIdx.x, p.z, p.gr, sh.NX, sh.NY, p.lambda_tsv, p.part + p.L.tsv_z, red);
}

__global__ void __launch_bounds__(MF_BLOCK) kw_prox(const WProb* __restrict__ W, WShared sh) {
  const WProb& p = W[blockIdx.y];
  if (p.S->gate_trial) return;
  __shared__ double red[16 * kProxW];
  dev_prox<double>(blockIdx.x, p.z, p.gr, p.xnew, p.d, p.box, sh.NX, sh.NY, p.S, p.part + p.L.prox, red);
}

__global__ void __launch_bounds__(MF_BLOCK) kw_control(const WProb* __restrict__ W, int* ndone) {
  const WProb& p = W[blockIdx.y];
  __shared__ double red[16 * 9];
  const int was_done = p.S->done;
  dev_control(p.S, p.cost, p.part, p.L, red);
  if (threadIdx.x == 0 && !was_done && p.S->done) atomicAdd(ndone, 1);
}

// the momentum step in both domains: blocks [0, n_elem) the pixels
// (z = a x + b xnew, x <- xnew on acceptance), the rest the models
__global__ void __launch_bounds__(MF_BLOCK) kw_momentum(const WProb* __restrict__ W, WShared sh, int n_elem) {
  const WProb& p = W[blockIdx.y];
  if ((int)blockIdx.x < n_elem)
    dev_momentum<double>(blockIdx.x, p.x, p.xnew, p.z, sh.NX * sh.NY, p.S);
  else
    vis_momentum(p, blockIdx.x - n_elem, 0);
}

__global__ void __launch_bounds__(MF_BLOCK) kw_stats(const WProb* __restrict__ W, WShared sh) {
  const WProb& p = W[blockIdx.y];
  __shared__ double red[64];
  dev_stats<double>(blockIdx.x, p.x, sh.NX, sh.NY, p.stat, red);
}

__global__ void __launch_bounds__(MF_BLOCK) kw_init_control(const WProb* __restrict__ W, int n_elem) {
  const WProb& p = W[blockIdx.y];
  __shared__ double red[16];
  dev_init_control(p.S, p.part, p.L, p.stat, n_elem, red);
}

// ---------------------------------------------------------------------------
// Toeplitz steps.

// the gradient at z without its TSV term, b - T z, and the partials of
// chi^2(z) - |W V|^2 = <z, T z> - 2 <z, b> (L.chi_z, one per element block)
__global__ void __launch_bounds__(MF_BLOCK) kwt_grad(const WProb* __restrict__ W, WShared sh) {
  const WProb& p = W[blockIdx.y];
  if (p.S->gate_prologue) return;
  __shared__ double red[16];
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (f < sh.NX * sh.NY) {
    const double b = p.bimg[f], tz = p.tz[f], z = p.z[f];
    p.gr[f] = b - tz;
    v[0] = z * (tz - 2.0 * b);
  }
  block_sum<1>(v, red);
  if (threadIdx.x == 0) p.part[p.L.chi_z + blockIdx.x] = v[0];
}

// row pass of the Toeplitz product of the start (init) or of a trial's step,
// two image rows per task (the pruned transform of toeplitz_body.h)
__global__ void __launch_bounds__(MF_BLOCK) kwt_rowfft(const WProb* __restrict__ W, WShared sh, int which) {
  const WProb& p = W[blockIdx.y];
  if (w_closed(p.S, which)) return;
  extern __shared__ double4 smem_d4[];
  const double* x = (which == kWInit) ? p.x : p.d;
  const int NY = sh.NY;
  auto load = [&](int i, int j) -> double { return x[(size_t)i * NY + j]; };
  st_rowfft_fwd_h<double, decltype(load), false>(blockIdx.x, sh.NX, NY, sh.Pxh, sh.twxh64, sh.twy64, p.T64,
                                                 reinterpret_cast<double2*>(smem_d4), load);
}

// column pass of the Toeplitz product of the start (init) or of a trial's
// step: task ky < NY (row 0 carrying the Nyquist row), the output in Gf64
__global__ void __launch_bounds__(MF_BLOCK) kwt_colconv(const WProb* __restrict__ W, WShared sh, int which) {
  const WProb& p = W[blockIdx.y];
  if (w_closed(p.S, which)) return;
  extern __shared__ double4 smem_d4[];
  st_colconv1<double>(blockIdx.x, p.T64, p.kh, p.Gf64, sh.NX, sh.NY + 1, sh.Lx, sh.NX / 2, sh.Pxh, sh.twxh64,
                      sh.twx64, reinterpret_cast<double2*>(smem_d4));
}

// c2r of the column pass's output, two image rows per task:
//   init:  T x = T z = the result; partials of <x, T x> - 2 <x, b> (L.chi_x)
//   trial: T xnew = T z + T d; partials of <xnew, T xnew> - 2 <xnew, b>
//          (L.chi_x) and of <d, T d> = |W A d|^2 (L.lip)
__global__ void __launch_bounds__(MF_BLOCK) kwt_c2r(const WProb* __restrict__ W, WShared sh, int which) {
  const WProb& p = W[blockIdx.y];
  if (w_closed(p.S, which)) return;
  extern __shared__ double4 smem_d4[];
  __shared__ double red[32];
  const int NY = sh.NY;
  double v[2] = {0.0, 0.0};
  const bool init = which == kWInit;
  auto store = [&](int i, int j, double td) {
    const size_t f = (size_t)i * NY + j;
    if (init) {
      p.tx[f] = td;
      p.tz[f] = td;
      v[0] += p.x[f] * (td - 2.0 * p.bimg[f]);
    } else {
      const double xv = p.xnew[f], tn = p.tz[f] + td;
      p.txn[f] = tn;
      v[0] += xv * (tn - 2.0 * p.bimg[f]);
      v[1] += p.d[f] * td;
    }
  };
  st_rowfft_c2r_h<double>(blockIdx.x, p.Gf64, sh.NX, NY, sh.Pxh, sh.twxh64, sh.twy64,
                          reinterpret_cast<double2*>(smem_d4), store);
  block_sum<2>(v, red);
  if (threadIdx.x == 0) {
    p.part[p.L.chi_x + blockIdx.x] = v[0];
    if (!init) p.part[p.L.lip + blockIdx.x] = v[1];
  }
}

// the momentum step in both images: z = a x + b xnew, T z = a T x + b T xnew;
// x <- xnew and T x <- T xnew on acceptance
__global__ void __launch_bounds__(MF_BLOCK) kwt_momentum(const WProb* __restrict__ W, WShared sh) {
  const WProb& p = W[blockIdx.y];
  const DevState* S = p.S;
  if (S->gate_momentum) return;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= sh.NX * sh.NY) return;
  const double a = S->mom_a, b = S->mom_b;
  const double xf = p.x[f], xn = p.xnew[f], tf = p.tx[f], tn = p.txn[f];
  p.z[f] = xf * a + b * xn;
  p.tz[f] = tf * a + b * tn;
  if (S->accept_x) {
    p.x[f] = xn;
    p.tx[f] = tn;
  }
}

// ---------------------------------------------------------------------------
// Host side.

static int nseq_for(int L, size_t csize, int nseq_total) {
  int n = (int)(65536 / (2 * (size_t)L * csize));
  if (n < 1) n = 1;
  int want = nseq_total / 256;
  if (want < 1) want = 1;
  if (n > want) n = want;
  if (n > 8) n = 8;
  return n;
}

struct WLaunch {
  dim3 rowf, colf, deg, grid, colinv, c2r, elem, one, vis, conv, c2r64;
  size_t sh_rowf, sh_colf, sh_grid, sh_colinv, sh_c2r, sh_conv, sh_c2r64;
};

static void launch_forward(const WLaunch& L, const WProb* W, const WShared& sh, int which, int MSP,
                           hipStream_t st) {
  hipLaunchKernelGGL(kw_rowfft64, L.rowf, dim3(MF_BLOCK), L.sh_rowf, st, W, sh, which);
  hipLaunchKernelGGL(kw_colfft64, L.colf, dim3(MF_BLOCK), L.sh_colf, st, W, sh, which);
  if (MSP == 6)
    hipLaunchKernelGGL(kw_degrid64<6>, L.deg, dim3(MF_BLOCK), 0, st, W, sh, which);
  else
    hipLaunchKernelGGL(kw_degrid64<12>, L.deg, dim3(MF_BLOCK), 0, st, W, sh, which);
}

// one trial step of every problem (the stream engine's step, with the
// residual at z from the models)
static void launch_step(const WLaunch& L, const WProb* W, const WShared& sh, int MSP, hipStream_t st) {
  const int cpt = (sh.TX * sh.TY) / MF_BLOCK;
  if (cpt == 1)
    hipLaunchKernelGGL(kw_grid32<1>, L.grid, dim3(MF_BLOCK), L.sh_grid, st, W, sh);
  else if (cpt == 2)
    hipLaunchKernelGGL(kw_grid32<2>, L.grid, dim3(MF_BLOCK), L.sh_grid, st, W, sh);
  else
    hipLaunchKernelGGL(kw_grid32<4>, L.grid, dim3(MF_BLOCK), L.sh_grid, st, W, sh);
  hipLaunchKernelGGL(kw_colinv32, L.colinv, dim3(MF_BLOCK), L.sh_colinv, st, W, sh);
  hipLaunchKernelGGL(kw_c2r32, L.c2r, dim3(MF_BLOCK), L.sh_c2r, st, W, sh);
  hipLaunchKernelGGL(kw_tsv_grad, L.elem, dim3(MF_BLOCK), 0, st, W, sh);
  hipLaunchKernelGGL(kw_prox, L.elem, dim3(MF_BLOCK), 0, st, W, sh);
  launch_forward(L, W, sh, kWTrial, MSP, st);
  hipLaunchKernelGGL(kw_control, L.one, dim3(MF_BLOCK), 0, st, W, sh.ndone);
  hipLaunchKernelGGL(kw_momentum, dim3(L.elem.x + L.vis.x, L.elem.y), dim3(MF_BLOCK), 0, st, W, sh, (int)L.elem.x);
}

// one Toeplitz trial step of every problem
static void launch_step_toep(const WLaunch& L, const WProb* W, const WShared& sh, hipStream_t st) {
  hipLaunchKernelGGL(kwt_grad, L.elem, dim3(MF_BLOCK), 0, st, W, sh);
  hipLaunchKernelGGL(kw_tsv_grad, L.elem, dim3(MF_BLOCK), 0, st, W, sh);
  hipLaunchKernelGGL(kw_prox, L.elem, dim3(MF_BLOCK), 0, st, W, sh);
  hipLaunchKernelGGL(kwt_rowfft, L.c2r64, dim3(MF_BLOCK), L.sh_c2r64, st, W, sh, (int)kWTrial);
  hipLaunchKernelGGL(kwt_colconv, L.conv, dim3(MF_BLOCK), L.sh_conv, st, W, sh, (int)kWTrial);
  hipLaunchKernelGGL(kwt_c2r, L.c2r64, dim3(MF_BLOCK), L.sh_c2r64, st, W, sh, (int)kWTrial);
  hipLaunchKernelGGL(kw_control, L.one, dim3(MF_BLOCK), 0, st, W, sh.ndone);
  hipLaunchKernelGGL(kwt_momentum, L.elem, dim3(MF_BLOCK), 0, st, W, sh);
}

int solve_nufft_wide(Context& C, const std::vector<NufftState*>& sets, const std::vector<RunParams>& P,
                     std::vector<RunStats>& out, float* kernel_ms, const std::vector<const ToepOut*>* ops) {
  const int B = (int)P.size();
  if (B <= 0 || (int)sets.size() != B) return MFISTA_E_ARG;
  const bool toep = ops != nullptr;
  if (toep && ((int)ops->size() != B || P[0].NX != P[0].NY || P[0].NX % 2)) return MFISTA_E_ARG;
  NufftState& N0 = *sets[0];
  const int NX = P[0].NX, NY = P[0].NY, n = NX * NY, MSP = N0.MSP;
  for (int b = 0; b < B; b++) {
    const NufftState& N = *sets[b];
    if (P[b].NX != NX || P[b].NY != NY || N.MSP != MSP || N.geom.TX != N0.geom.TX || N.geom.TY != N0.geom.TY ||
        P[b].tv_mode)
      return MFISTA_E_ARG;
  }
  hipStream_t st = C.stream;
  const GridGeom& g0 = N0.geom;
  const int nky = NY + 1;
  WShared sh;
  memset(&sh, 0, sizeof(sh));
  sh.Px = N0.Px;
  sh.Py = N0.Py;
  sh.twx64 = (const double2*)N0.twxd.p;
  sh.twy64 = (const double2*)N0.twyd.p;
  sh.twx32 = (const float2*)N0.twxf.p;
  sh.twy32 = (const float2*)N0.twyf.p;
  sh.e4x64 = (const double*)N0.e4xd.p;
  sh.e4y64 = (const double*)N0.e4yd.p;
  sh.e4x32 = (const float*)N0.e4xf.p;
  sh.e4y32 = (const float*)N0.e4yf.p;
  sh.NX = NX;
  sh.NY = NY;
  sh.Lx = g0.Lx;
  sh.Ly = g0.Ly;
  sh.MSP = MSP;
  sh.TX = g0.TX;
  sh.TY = g0.TY;
  sh.nseq_r64 = nseq_for(g0.Ly, sizeof(double2), (NX + 1) / 2);
  sh.nseq_c64 = nseq_for(g0.Lx, sizeof(double2), nky);
  sh.nseq_ci32 = nseq_for(g0.Lx, sizeof(float2), nky);
  sh.nseq_c2r32 = nseq_for(g0.Ly, sizeof(float2), (NX + 1) / 2);
  sh.adj_scale = (float)(1.0 / (8.0 * (double)NX * NY));
  if (toep) {  // every problem has NX = NY = the batch's size: the first set's plan serves all
    sh.Pxh = (*ops)[0]->Pxh;
    sh.twxh64 = (const double2*)(*ops)[0]->twxh64.p;
  }
  // per-problem workspaces and descriptors
  if ((int)C.wwork.size() < B) C.wwork.resize(B);
  std::vector<WProb> hw(B);
  const int n_elem = elem_blocks(n);
  const size_t rb = (size_t)n * 8;

  for (int b = 0; b < B; b++) {
    NufftState& N = *sets[b];
    WideWork& Wk = C.wwork[b];
    const RunParams& R = P[b];
    // partials of a trial's chi^2 and |W A d|^2: one per degrid task, or per
    // c2r task (two image rows) in Toeplitz steps
    const int ndeg = toep ? (NX + 1) / 2 : degrid_blocks(N.M, MSP);
    PartialLayout L;
    L.n_degrid = ndeg;
    L.n_elem = n_elem;
    L.chi_z = 0;
    L.chi_x = ndeg;
    L.lip = 2 * ndeg;
    L.tsv_z = 3 * ndeg;
    L.prox = L.tsv_z + n_elem;
    L.total = L.prox + kProxW * n_elem;
    // chi^2 at z comes from the visibility-domain momentum, one partial per
    // 256 visibilities (Toeplitz steps: from the gradient kernel, one per
    // element block), after the other regions
    L.n_chiz = toep ? n_elem : (N.M + MF_BLOCK - 1) / MF_BLOCK;
    L.chi_z = L.total;
    L.total += L.n_chiz;
    if (Wk.x.ensure(rb) || Wk.z.ensure(rb) || Wk.xnew.ensure(rb) || Wk.gr.ensure(rb) || Wk.d.ensure(rb) ||
        Wk.T64.ensure((size_t)nky * NX * sizeof(double2)) ||
        Wk.Gf64.ensure(((size_t)NY + MSP) * g0.Lx * sizeof(double2)) || Wk.part.ensure((size_t)L.total * 8) ||
        Wk.stat.ensure((size_t)4 * n_elem * 8) || Wk.state.ensure(sizeof(DevState)) ||
        Wk.cost.ensure((size_t)(R.maxiter > 0 ? R.maxiter : 1) * 8))
      return MFISTA_E_HIP;
    if (toep) {
      if (Wk.tvec.ensure(3 * rb)) return MFISTA_E_HIP;
    } else {
      if (Wk.mod.ensure((size_t)3 * N.M * sizeof(double2)) || Wk.rz32.ensure((size_t)N.M * sizeof(float2)) ||
          Wk.Hg.ensure((size_t)nky * g0.Lx * sizeof(float2)) || Wk.U.ensure((size_t)NX * nky * sizeof(float2)))
        return MFISTA_E_HIP;
      MF_CHECK(hipMemsetAsync(Wk.Hg.p, 0, (size_t)nky * g0.Lx * sizeof(float2), st));
    }
    MF_CHECK(hipMemcpyAsync(Wk.x.p, R.xinit, rb, hipMemcpyDeviceToDevice, st));
    MF_CHECK(hipMemcpyAsync(Wk.z.p, R.xinit, rb, hipMemcpyDeviceToDevice, st));
    DevState h;
    memset(&h, 0, sizeof(h));
    h.c = R.cinit;
    h.lambda_l1 = R.lambda_l1;
    h.lambda_tsv = R.lambda_tsv;
    h.eps_abs = R.eps_abs;
    h.rel_tol = R.rel_tol;
    h.eta_first = R.eta_first;
    h.maxiter = R.maxiter;
    h.nonneg = R.nonneg;
    h.lip_mode = 1;
    h.target_F = R.target_F;
    h.chi_add = toep ? (*ops)[b]->wv2 : 0.0;
    MF_CHECK(hipMemcpyAsync(Wk.state.p, &h, sizeof(h), hipMemcpyHostToDevice, st));
    WProb& w = hw[b];
    memset(&w, 0, sizeof(w));
    w.g = N.geom;
    w.pd = N.pd();
    w.pf = N.pf();
    w.chunks = (const int4*)N.chunks.p;
    w.ent = (const unsigned*)N.ent.p;
    w.nchunks = N.nchunks;
    w.n_degrid = ndeg;
    w.L = L;
    w.x = (double*)Wk.x.p;
    w.z = (double*)Wk.z.p;
    w.xnew = (double*)Wk.xnew.p;
    w.gr = (double*)Wk.gr.p;
    w.d = (double*)Wk.d.p;
    w.T64 = (double2*)Wk.T64.p;
    w.Gf64 = (double2*)Wk.Gf64.p;
    w.mx = (double2*)Wk.mod.p;
    w.mxn = w.mx + N.M;
    w.mz = w.mxn + N.M;
    w.M = N.M;
    w.rz32 = (float2*)Wk.rz32.p;
    w.Hg = (float2*)Wk.Hg.p;
    w.U = (float2*)Wk.U.p;
    w.part = (double*)Wk.part.p;
    w.stat = (double*)Wk.stat.p;
    w.cost = (double*)Wk.cost.p;
    w.S = (DevState*)Wk.state.p;
    w.box = R.box_flag ? R.box : nullptr;
    w.lambda_tsv = R.lambda_tsv;
    if (toep) {
      w.kh = (const double*)(*ops)[b]->kh.p;
      w.bimg = (const double*)(*ops)[b]->bimg.p;
      w.tx = (double*)Wk.tvec.p;
      w.txn = w.tx + n;
      w.tz = w.txn + n;
      w.mx = w.mxn = w.mz = nullptr;
      w.rz32 = nullptr;
      w.Hg = w.U = nullptr;
    }
  }
  if (C.wdesc.ensure((size_t)B * sizeof(WProb)) || C.wdone.ensure(sizeof(int))) return MFISTA_E_HIP;
  MF_CHECK(hipMemcpyAsync(C.wdesc.p, hw.data(), (size_t)B * sizeof(WProb), hipMemcpyHostToDevice, st));
  MF_CHECK(hipMemsetAsync(C.wdone.p, 0, sizeof(int), st));
  sh.ndone = (int*)C.wdone.p;
  const WProb* W = (const WProb*)C.wdesc.p;
  // launch shapes for the problems hw[idx[0..n)] (descriptors in that order)
  std::vector<int> act(B);
  for (int b = 0; b < B; b++) act[b] = b;
  auto shape = [&](const std::vector<int>& idx) {
    const int n = (int)idx.size();
    int mdeg = 1, mch = 1, mvb = 1;
    for (int b : idx) {
      mdeg = std::max(mdeg, hw[b].n_degrid);
      mch = std::max(mch, hw[b].nchunks);
      mvb = std::max(mvb, hw[b].L.n_chiz);
    }
    WLaunch L;
    L.rowf = dim3((NX + 2 * sh.nseq_r64 - 1) / (2 * sh.nseq_r64), n);
    L.sh_rowf = 2 * (size_t)sh.nseq_r64 * g0.Ly * sizeof(double2);
    L.colf = dim3((nky + sh.nseq_c64 - 1) / sh.nseq_c64, n);
    L.sh_colf = 2 * (size_t)sh.nseq_c64 * g0.Lx * sizeof(double2);
    L.deg = dim3(mdeg, n);
    L.grid = dim3((mch + kGB - 1) / kGB, n);
    L.sh_grid = grid_sep_lds_bytes(g0.TX, g0.TY);
    L.colinv = dim3((nky + sh.nseq_ci32 - 1) / sh.nseq_ci32, n);
    L.sh_colinv = 2 * (size_t)sh.nseq_ci32 * g0.Lx * sizeof(float2);
    L.c2r = dim3((NX + 2 * sh.nseq_c2r32 - 1) / (2 * sh.nseq_c2r32), n);
    L.sh_c2r = 2 * (size_t)sh.nseq_c2r32 * g0.Ly * sizeof(float2);
    L.elem = dim3(n_elem, n);
    L.one = dim3(1, n);
    L.vis = dim3(mvb, n);
    L.conv = dim3(NY, n);  // st_colconv1: NY tasks (row 0 carries the Nyquist row)
    L.sh_conv = 2 * (size_t)g0.Lx * sizeof(double2);
    L.c2r64 = dim3((NX + 1) / 2, n);
    L.sh_c2r64 = 2 * (size_t)g0.Ly * sizeof(double2);
    return L;
  };
  WLaunch L = shape(act);
  // start: the model of xinit, its cost (chi^2, |x|_1, TSV), z = x in both
  // domains, and the initial control state
  if (toep) {  // T x0 (= T z), its chi^2 partials
    hipLaunchKernelGGL(kwt_rowfft, L.c2r64, dim3(MF_BLOCK), L.sh_c2r64, st, W, sh, (int)kWInit);
    hipLaunchKernelGGL(kwt_colconv, L.conv, dim3(MF_BLOCK), L.sh_conv, st, W, sh, (int)kWInit);
    hipLaunchKernelGGL(kwt_c2r, L.c2r64, dim3(MF_BLOCK), L.sh_c2r64, st, W, sh, (int)kWInit);
  } else {
    launch_forward(L, W, sh, kWInit, MSP, st);
    hipLaunchKernelGGL(kw_vis_momentum, L.vis, dim3(MF_BLOCK), 0, st, W, 1);
  }
  hipLaunchKernelGGL(kw_stats, L.elem, dim3(MF_BLOCK), 0, st, W, sh);
  hipLaunchKernelGGL(kw_init_control, L.one, dim3(MF_BLOCK), 0, st, W, n_elem);
  MF_CHECK(hipGetLastError());
  // one step captured as a graph; re-captured for the problems still running
  // whenever they have become half as many (a batch waits for its slowest
  // problem, and finished problems would otherwise keep their workgroups in
  // every launch)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  auto capture = [&]() -> int {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    MF_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    if (toep)
      launch_step_toep(L, W, sh, st);
    else
      launch_step(L, W, sh, MSP, st);
    if (hipStreamEndCapture(st, &graph) != hipSuccess) return MFISTA_E_HIP;
    if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) return MFISTA_E_HIP;
    return 0;
  };
  int rc = capture();
  int maxit = 1;
  for (const RunParams& R : P) maxit = std::max(maxit, R.maxiter);
  const long cap = (long)maxit * 64 + 100000;
  const int K = 16;  // steps between reads of the finished-problem counter
  int* hflag = C.hflags;
  int pending = -1;
  long step = 0;
  for (hipEvent_t& e : C.ev_t)
    if (!e) MF_CHECK(hipEventCreate(&e));
  MF_CHECK(hipEventRecord(C.ev_t[0], st));
  for (; rc == 0; step++) {
    if (hipGraphLaunch(exec, st) != hipSuccess) {
      rc = MFISTA_E_HIP;
      break;
    }
    if ((step + 1) % K == 0) {
      const int slot = (int)((step / K) & 1);
      MF_CHECK(hipMemcpyAsync(&hflag[slot], C.wdone.p, sizeof(int), hipMemcpyDeviceToHost, st));
      MF_CHECK(hipEventRecord(C.ev[slot], st));
      if (pending >= 0) {
        MF_CHECK(hipEventSynchronize(C.ev[pending]));
        const int ndone = hflag[pending];
        if (ndone >= B) break;
        if ((int)act.size() >= 4 && 2 * (B - ndone) <= (int)act.size()) {
          MF_CHECK(hipStreamSynchronize(st));
          std::vector<int> keep;
          std::vector<WProb> sub;
          for (int b : act) {
            int done = 0;
            MF_CHECK(hipMemcpy(&done, &((DevState*)C.wwork[b].state.p)->done, sizeof(int), hipMemcpyDeviceToHost));
            if (!done) {
              keep.push_back(b);
              sub.push_back(hw[b]);
            }
          }
          if (keep.empty()) break;
          act = keep;
          MF_CHECK(hipMemcpyAsync(C.wdesc.p, sub.data(), sub.size() * sizeof(WProb), hipMemcpyHostToDevice, st));
          L = shape(act);
          if ((rc = capture())) break;
          pending = -1;
          continue;
        }
      }
      pending = slot;
    }
    if (step > cap) {
      fprintf(stderr, "[mfista_amd] wide engine: solver did not terminate\n");
      rc = MFISTA_E_ARG;
      break;
    }
  }
  MF_CHECK(hipEventRecord(C.ev_t[1], st));
  MF_CHECK(hipStreamSynchronize(st));
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc) return rc;
  float ms = 0;
  MF_CHECK(hipEventElapsedTime(&ms, C.ev_t[0], C.ev_t[1]));
  if (kernel_ms) *kernel_ms = ms;
  out.assign(B, RunStats());
  for (int b = 0; b < B; b++) {
    DevState h;
    MF_CHECK(hipMemcpy(&h, C.wwork[b].state.p, sizeof(h), hipMemcpyDeviceToHost));
    RunStats& o = out[b];
    o.ITER = h.ITER;
    o.c = h.c;
    o.max_trials = h.max_trials;
    o.total_trials = h.total_trials;
    o.status = h.status;
    o.steps = step + 1;
    o.rounds = h.total_trials;
    o.kernel_ms = ms;
    MF_CHECK(hipMemcpyAsync(P[b].xout, C.wwork[b].x.p, rb, hipMemcpyDeviceToDevice, st));
  }
  MF_CHECK(hipStreamSynchronize(st));
  C.last_steps = step + 1;
  return 0;
}

}  // namespace mf
