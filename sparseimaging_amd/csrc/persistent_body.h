// persistent_body.h -- device code of the persistent NUFFT MFISTA solver.
//
// A 256^2 solve moves a few MB per iteration, a few microseconds of HBM/L2
// time, so a stream of ~13 small kernels per trial is bound by launch and
// dispatch latency.  Here one cooperative launch (all workgroups co-resident)
// runs the reference's loop (c/mfista_nufft_lib.c:557-787, L1+TSV; :308-553,
// L1+TV) end to end.  Stages of stages.h are dealt round-robin to the
// workgroups and separated by a grid barrier; the control step needs no
// barrier of its own: every workgroup reduces the same FP64 partial sums in
// the same order and reaches the same decision.
//
// Visibility-domain momentum.  The operator is linear, so the weighted model
// of the momentum point, W A z with z = a x + b xnew, is a W A x + b W A xnew.
// The solver keeps three FP64 model vectors (of x, of the last xnew, of z) and
// each trial computes W A d for its step d = xnew - z with an FP64 forward
// transform, so W A xnew = W A z + W A d.  The residual at z therefore needs
// no forward transform at all; z is held in FP64 so that the models stay the
// exact images of the stored iterates (no drift).  Per iteration with b
// trials (L1+TSV): grid | inverse column pass | c2r + gradient + prox + FP64
// row pass of d | FP64 column pass | FP64 degrid, then 3 stages per further
// trial: 5 + 3(b-1) grid barriers.  L1+TV keeps the c2r stage separate (the
// FGP iterations need the whole gradient) and adds one barrier per FGP step.
//
// The kernels are templates over the preset (MSPC), the placement of the
// twiddle and E4 tables (TAB) and the gridding tile (CPT cells per thread:
// 1 = 16 x 16, 2 = 32 x 16, 4 = 32 x 32 cells).  Each tile size is
// instantiated in its own translation unit (persistent_tile*.hip), so that
// the large kernel bodies compile in parallel; the host side
// (persistent.hip) reaches them through persistent_kernel<CPT>().
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "deal.h"
#include "driver.h"
#include "stages.h"

namespace mf {

struct PArgs {
  GridGeom g;
  FFTPlan Px, Py;
  VisParams<float> pf;
  VisParams<double> pd;
  const float2 *twxf, *twyf;
  const double2 *twxd, *twyd;
  const float *e4xf, *e4yf;
  const double *e4xd, *e4yd;
  float2 *Hg, *U;
  double2 *T64, *Gf64;    // FP64 forward transform of the step (first trial of a round)
  double2 *T64b, *Gf64b;  // the same for the second trial of a paired round
  double2* vm[4];         // weighted model vectors W A (.), roles rotate on acceptance
  const int4* chunks;
  int nchunks;
  const int* granges;  // chunk range [granges[b], granges[b+1]) of workgroup b
  const unsigned* ent;
  double *x, *z;  // iterates in FP64 (FP32 iterates stagnate near a tight tolerance)
  double* xn[2];  // xnew of the first and second trial of a round
  float* gr;
  const float* box;
  float *rA, *sA, *rB, *sB, *pp, *qq;
  const double* fgp_c;  // c1[k] = fgp_c[2k], c2[k] = fgp_c[2k+1]
  int fgp_iter, tv_mode, maxiter, nonneg;
  double lambda_l1, lambda_tsv, lambda_tv, eps_abs, rel_tol, eta_first, cinit, adj_scale;
  double* part;  // kPW doubles per workgroup (see kP* below)
  int n_deg, n_row, n_ky, n_elem;
  int lds_tables;  // twiddle and E4 tables staged in LDS
  int nblk;        // workgroups solving this problem
  double* cost;
  DevState* S;
  unsigned long long* bar;  // barrier words, see BarCtx
  int fgp_tw;               // FGP block shape: 0 = one pixel per thread along the flat image,
                            // else tiles of fgp_tw x (256 / fgp_tw) pixels per workgroup
  int hier_barrier;         // XCD-hierarchical grid barrier
  int pair;                 // L1+TSV: evaluate two backtracking trials per round
  // Coupled movie (frames of one solve, matlab/MFISTA_movie_eht.m): every
  // frame is one PArgs of a batch launch; all frames share one barrier, one
  // partial-sum table and so one control step.  ngb = 0: an ordinary solve.
  unsigned long long* trace;  // per-barrier clocks (nullptr: no trace), see BarCtx
  int* ttag;
  int tcap;
  int ngb;                   // workgroups of all frames (barrier, partial rows)
  int gb0;                   // partial row of this frame's workgroup 0
  double lambda_t;           // weight of d_X = sum_t w_t |x_t - x_{t+1}|^2
  double wt_prev, wt_next;   // w_{t-1}, w_t (0 at the first / last frame)
  const double *z_prev, *z_next;  // the neighbouring frames' z (nullptr at the ends)
  const double* x_next;           // the next frame's x (initial cost)
  const double* xn_next[2];       // the next frame's trial images
  // Toeplitz loop (toeplitz_body.h): Kh[ky][k] the real spectrum of the
  // data term's normal operator, b = Re A^H W^2 V, the operator applied to
  // x, xnew (two trials) and z (roles as vm[]), the column-pass outputs of
  // the two trials, |W V|^2
  const double* kh;
  const double* bimg;
  double* tv[4];
  double2 *u64a, *u64b;
  double wv2;
  double* xb[4];  // image buffers whose roles rotate (BlockState::xr)
  FFTPlan Pxh;           // the NX-point plan of the pruned row and column transforms
  const double2* twxh;   // and its twiddles
};

// Grid-wide barriers for a cooperative launch.  bar[] layout (64-bit words):
//   [0] arrivals of the flat barrier, or of the XCD leaders
//   [1] error flag (a bounded wait that runs out sets it; every wait checks it)
// Every signal word is written with an atomic read-modify-write: those are
// performed at the device coherence point, where the polling loads (relaxed,
// agent scope) read; a plain store would stay in the writer's L2.
//   [2] arrivals of the set-up barrier, [3] XCD id out of range
//   [4..7] the first failed wait: word, target, value seen, workgroup | polls << 20
//   XCD x: [kXBase + kXStride x + 0] local arrivals, [+8] members, [+16] release epoch
// (each on its own 64-byte line).
constexpr int kXBase = 32, kXStride = 32;
constexpr size_t kBarBytes = (size_t)(kXBase + 8 * kXStride) * 8;
constexpr long kSpinLimit = 1L << 26;

struct BarCtx {
  unsigned long long* bar;
  unsigned nblk;   // workgroups in the launch
  unsigned nx;     // workgroups on this workgroup's XCD
  unsigned ngrp;   // XCDs with workgroups
  int xcc;         // this workgroup's XCD
  int leader;      // first workgroup of its XCD
  int hier;        // 1: XCD-hierarchical barrier, 0: flat
  // tracing (MFISTA_PERSISTENT_TRACE): per barrier and workgroup, the clock at
  // arrival and at release; the stage tag of each barrier from workgroup 0
  unsigned long long* trace;
  int* ttag;
  int tcap, trow, tnrow;
};

// bounded wait until bar[i] >= target; false (and the error flag set) on
// time-out.  RMW = true polls with atomic read-modify-writes (fetch_add 0),
// which are performed at the coherence point: measured on gfx950, relaxed
// atomic loads of a word that one workgroup increments are seen by the other
// workgroups of its XCD only after ~20 us, RMW polls within ~1 us.
template <bool RMW>
__device__ inline bool wait_ge(unsigned long long* bar, int i, unsigned long long target) {
  long spins = 0;
  while (true) {
    const unsigned long long v =
        RMW ? __hip_atomic_fetch_add(&bar[i], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
            : __hip_atomic_load(&bar[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v >= target) return true;
    __builtin_amdgcn_s_sleep(1);
    if ((++spins & 1023) == 0 &&
        (__hip_atomic_fetch_add(&bar[1], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) || spins > kSpinLimit)) {
      // the first workgroup to give up records what it waited for
      if (__hip_atomic_fetch_or(&bar[1], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        __hip_atomic_fetch_or(&bar[4], (unsigned long long)i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_or(&bar[5], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_or(&bar[6], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_or(&bar[7], (unsigned long long)blockIdx.x | ((unsigned long long)spins << 20),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      return false;
    }
  }
}

// Flat barrier on counter bar[ci]: one agent-scope release fence (L2
// write-back) before arriving and one acquire fence (L2 invalidate) after the
// wait, per workgroup.  The k-th barrier is passed once the monotonic counter
// reaches k * nblk, so no reset is needed.
__device__ inline bool flat_sync(unsigned long long* bar, int ci, unsigned nblk, unsigned long long epoch) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __hip_atomic_fetch_add(&bar[ci], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const bool ok = wait_ge<false>(bar, ci, epoch * (unsigned long long)nblk);
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return ok;
}

// XCD-hierarchical barrier.  The L2 is per XCD and the L1 is write-through,
// so once every wave of the XCD's workgroups has its stores acknowledged
// (s_waitcnt vmcnt(0)) the data sits in that XCD's L2.  Only the XCD leader
// then writes the L2 back (release fence) and arrives at the global counter;
// after all leaders have arrived it releases its XCD, and every workgroup
// takes an acquire fence.  One L2 write-back per XCD instead of one per
// workgroup, and 8-way instead of nblk-way contention on the global counter.
__device__ inline bool xcd_sync(const BarCtx& B, unsigned long long epoch) {
  unsigned long long* bar = GP(B.bar);
  const int xb = kXBase + kXStride * B.xcc;
  __hip_atomic_fetch_add(&bar[xb], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (B.leader) {
    if (!wait_ge<true>(bar, xb, epoch * (unsigned long long)B.nx)) return false;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(&bar[0], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (!wait_ge<true>(bar, 0, epoch * (unsigned long long)B.ngrp)) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    __hip_atomic_fetch_add(&bar[xb + 16], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // = epoch
    return true;
  }
  if (!wait_ge<true>(bar, xb + 16, epoch)) return false;
  // agent-scope acquire: a work-group-scope invalidate does not reach the L1
  // of a non-split work-group, which could then serve stale lines
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// tag: the stage that ends at this barrier (for the trace)
__device__ inline bool grid_sync(const BarCtx& B, unsigned long long& epoch, int tag = 0) {
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  __shared__ int fail;
  epoch += 1;
  if (threadIdx.x == 0) {
    const unsigned long long ta = wall_clock64();
    fail = (B.hier ? xcd_sync(B, epoch) : flat_sync(GP(B.bar), 0, B.nblk, epoch)) ? 0 : 1;
    if (B.trace && epoch <= (unsigned long long)B.tcap) {
      const unsigned long long tr = wall_clock64();
      unsigned long long* t = B.trace + ((epoch - 1) * (unsigned long long)B.tnrow + (unsigned long long)B.trow) * 2;
      t[0] = ta;
      t[1] = tr;
      if (B.trow == 0) B.ttag[epoch - 1] = tag;
    }
  }
  __syncthreads();
  return fail == 0;
}

// Set-up: every workgroup learns its XCD (hardware register XCC_ID), the XCD
// member counts and its leader; one flat barrier publishes the counts.  The
// hierarchical barrier is used only if every XCD id is in range.
__device__ inline bool bar_setup(BarCtx& B, unsigned long long* bar, unsigned nblk, int want_hier) {
  if (threadIdx.x == 0) {
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    B.bar = bar;
    B.nblk = nblk;
    B.xcc = (int)(v & 0xf);
    B.leader = 0;
    B.hier = 0;
    if (B.xcc < 8) {
      const unsigned long long idx = __hip_atomic_fetch_add(&bar[kXBase + kXStride * B.xcc + 8], 1ull,
                                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      B.leader = (idx == 0);
    } else {
      __hip_atomic_fetch_or(&bar[3], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // out of range
    }
  }
  __syncthreads();
  __shared__ int fail;
  if (threadIdx.x == 0) {
    fail = flat_sync(bar, 2, nblk, 1) ? 0 : 1;
    if (!fail) {
      unsigned ng = 0;
      for (int x = 0; x < 8; x++)
        ng += __hip_atomic_fetch_add(&bar[kXBase + kXStride * x + 8], 0ull, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT) ? 1u : 0u;
      B.ngrp = ng;
      B.nx = (B.xcc < 8) ? (unsigned)__hip_atomic_fetch_add(&bar[kXBase + kXStride * B.xcc + 8], 0ull,
                                                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         : 0u;
      const bool bad = __hip_atomic_fetch_add(&bar[3], 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
      B.hier = (want_hier && !bad) ? 1 : 0;
    }
  }
  __syncthreads();
  return fail == 0;
}

// Per-workgroup partial sums: kPW slots per workgroup, written by the stage
// that produces them and reduced by the control step of every workgroup, in
// the same fixed order everywhere.  Stored slot-major (part[slot * nrow + row])
// so that a wave reads a slot of all rows with coalesced loads: row-major, each
// lane of a load touched its own cache line and the reduction took ~13 us.
constexpr int kPW = 16;
// trial k of a round uses slots 6 k + (0 .. 5)
// (kPTx, kPTd: this frame's share of d_X(xnew) and d_X(d) in a coupled movie)
enum { kPD2 = 0, kPL1 = 1, kPReg = 2, kPTsvd = 3, kPChi = 4, kPLip = 5, kPTx = 6, kPTd = 7, kPTrial = 8 };
// init-phase slots reuse the same row: |x|_1, TSV(x), TV(x), d_X(x) share, chi^2(x)
enum { kPIl1 = 0, kPItsv = 1, kPItv = 2, kPIdx = 3, kPIchi = 4 };

template <int NV>
__device__ inline void part_reduce(const double* part, int nrow, const int (&slot)[NV], double (&out)[NV],
                                   double* red) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    if (q % nw != wid) continue;  // wave wid reduces slots wid, wid + nw, ...
    const double* col = part + (size_t)slot[q] * nrow;
    double s = 0.0;
#pragma unroll 8
    for (int i = lane; i < nrow; i += 64) s += gld(col + i);
    s = wave_sum(s);
    if (lane == 0) red[q] = s;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; q++) out[q] = red[q];
  __syncthreads();
}

struct BlockState {
  double c, mu, costtmp, l1_x, mom_a, mom_b;
  int iter, trial, accept, mom_pending, stop, ITER, max_trials, total_trials, status, prologue;
  int ix, in, iz;  // roles of the model vectors: W A x, W A xnew, W A z
  int in1;         // W A xnew of the second trial of a paired round
  int kacc;        // trial of the last round that was accepted (0 or 1): xn[kacc]
  int rounds;      // trial rounds evaluated
  int xr[4];       // Toeplitz loop: image buffers of x, the two trials' xnew, a spare
};

// regulariser sums at pixel f over its right and lower neighbours:
// r[0] += reg(xnew) (TSV, or isotropic TV in tv mode), r[1] += TSV(d),
// d = xnew - z
__device__ inline void reg_terms(const double* xn, const double* z, int f, int W, int H, bool tv,
                                 double r[2]) {
  const int c = f % W, rr = f / W;
  const double x0 = xn[f], d0 = x0 - z[f];
  double dh = 0, dv = 0;
  if (c < W - 1) {
    const double x1 = xn[f + 1];
    dh = x0 - x1;
    const double u = d0 - (x1 - z[f + 1]);
    r[1] += u * u;
  }
  if (rr < H - 1) {
    const double x1 = xn[f + W];
    dv = x0 - x1;
    const double u = d0 - (x1 - z[f + W]);
    r[1] += u * u;
  }
  if (tv) {
    if (c < W - 1 && rr < H - 1)
      r[0] += sqrt(dh * dh + dv * dv);
    else if (c < W - 1)
      r[0] += fabs(dh);
    else if (rr < H - 1)
      r[0] += fabs(dv);
  } else {
    r[0] += dh * dh + dv * dv;
  }
}


// this frame's share of the temporal terms of trial k (coupled movie): r[0] +=
// w_t (xnew_t - xnew_{t+1})^2 and r[1] += w_t (d_t - d_{t+1})^2 at pixel f,
// d = xnew - z (matlab/lib/d_X.m)
template <typename PA>
__device__ inline void temporal_terms(const PA& A, int k, int f, double r[2]) {
  const double a = A.xn[k][f], b = A.xn_next[k][f];
  const double e = a - b, d = (a - A.z[f]) - (b - A.z_next[f]);
  r[0] += A.wt_next * e * e;
  r[1] += A.wt_next * d * d;
}

// soft threshold (c/mfista_tools.c:81-141): v -> 0 on the box and for |v| <= eta
template <typename R>
__device__ inline R prox_p(R v, R eta, int nonneg, bool boxed) {
  if (boxed) return (R)0;
  if (nonneg) return (v < eta) ? (R)0 : v - eta;
  if (v >= eta) return v - eta;
  if (v <= -eta) return v + eta;
  return (R)0;
}

__device__ inline float fgp_proj(float v, float thr, int nonneg, bool boxed) {
  if (boxed) return 0.f;
  if (nonneg) return v < 0.f ? 0.f : v;
  if (v >= thr) return v - thr;
  if (v <= -thr) return v + thr;
  return 0.f;
}

__device__ inline double fgp_proj_d(double v, double thr, int nonneg, bool boxed) {
  if (boxed) return 0.0;
  if (nonneg) return v < 0.0 ? 0.0 : v;
  if (v >= thr) return v - thr;
  if (v <= -thr) return v + thr;
  return 0.0;
}

__device__ inline float lpq_f(const float* p, const float* q, int i, int j, int W, int H) {
  float v = 0.f;
  if (i < W - 1) v += p[(W - 1) * j + i];
  if (i > 0) v -= p[(W - 1) * j + i - 1];
  if (j < H - 1) v += q[W * j + i];
  if (j > 0) v -= q[W * (j - 1) + i];
  return v;
}

// Gridding of this workgroup's range [cb, ce) of the tile-ordered chunk list
// (ranges balanced by entry count on the host), owner-computes: thread t owns
// cells t, t + 256, ... of the current tile and gathers the deposits of every
// entry of the chunk onto them, so there are no LDS or intra-tile atomics and
// the sums stay in registers across consecutive chunks of one tile.  The
// entries of up to kGB chunks are loaded at once, one thread per entry, into
// LDS records (tap offsets relative to the tile, tap slopes, deposit value
// E1 w conj(r)); each record is then read by all lanes as a broadcast.  A tile
// is flushed when it changes: with plain stores
// when this workgroup processed all of the tile's chunks (Hg is zero there),
// otherwise with one global atomic per touched cell.  Deposits as st_grid
// (stages.h): visibility s at taps (l, l') with weight
// exp(l a_x - l^2 t3x + l' a_y - l'^2 t3y), conjugated at -k for mirror entries.
constexpr int kGB = MF_BLOCK / kGridChunk;

// One gridding entry as the gather loop reads it (48 bytes, three 16-byte
// LDS loads).  Cell (row, col) of the tile takes taps l = ox + sg col along x
// and l' = oy + sg row along y (sg = +1 direct, -1 mirror).  The loop works
// with u = l - 1/2 and u' = l' - 1/2: the cell is covered when |u| < MSP and
// |u'| < MSP, and its weight exp(l a_x - l^2 t3x + l' a_y - l'^2 t3y) is
// 2^((a2x u + a1x) u + (a2y u' + a1y) u' + a0), the coefficients holding
// log2(e) and the half-tap shift.
struct alignas(16) GridEnt {
  float ux, uy, sg, a0;    // u and u' of cell (0, 0), direction, constant term
  float a1x, a2x, a1y, a2y;
  float dx, dy, pad0, pad1;  // deposit E1 w r (conjugated for mirror entries)
};

template <int CPT, typename Resid>
__device__ inline void grid_chunks(const GridGeom& g, const VisParams<float>& vp, Resid resid,
                                   const int4* __restrict__ chunks, int cb, int ce,
                                   const unsigned* __restrict__ ent, float2* __restrict__ Hg, char* lds) {
  const int TX = g.TX, TY = g.TY;
  GridEnt* er = reinterpret_cast<GridEnt*>(lds);
  int4* cmeta = reinterpret_cast<int4*>(er + MF_BLOCK);  // kGB chunk records
  const int NX = g.NX;
  const float fmsp = (float)g.MSP;
  const double t3x = g.t3x, t3y = g.t3y;
  constexpr double kL2E = 1.4426950408889634074;  // log2(e)
  // the cells this thread owns: (rr[c], cc[c]) within the tile
  int rr[CPT], cc[CPT];
  float frr[CPT], fcc[CPT];
#pragma unroll
  for (int c = 0; c < CPT; c++) {
    const int cell = threadIdx.x + c * MF_BLOCK;
    rr[c] = cell / TX;
    cc[c] = cell - rr[c] * TX;
    frr[c] = (float)rr[c];
    fcc[c] = (float)cc[c];
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
  for (int c0 = cb; c0 < ce; c0 += kGB) {
    const int nc = min(kGB, ce - c0);
    __syncthreads();  // the previous batch's entries are consumed
    {
      const int q = threadIdx.x, j = q / kGridChunk;
      if (j < nc) {
        const int4 ch = chunks[c0 + j];
        if ((q % kGridChunk) == 0) cmeta[j] = ch;
        const int e = ch.y + (q % kGridChunk);
        if (e < ch.z) {
          const unsigned en = ent[e];
          const int sv = (int)(en & 0x7fffffffu);
          const bool mir = (en >> 31) != 0;
          const int2 m = GP(vp.mxy)[sv];
          const float2 a = GP(vp.a)[sv];
          const float ew = GP(vp.e1)[sv] * GP(vp.w)[sv];
          const float2 r = resid(sv);
          const int tr = ch.x / g.ntx, tc = ch.x - tr * g.ntx;
          const int trb = tr * TY, tcb = tc * TX;
          // tap offsets of cell (row, col): direct l' = row - my, l = col - mx - NX;
          // mirror l' = -row - my, l = NX - mx - col
          const int oy = mir ? -(trb + m.y) : trb - m.y;
          const int ox = mir ? NX - m.x - tcb : tcb - m.x - NX;
          // l a - l^2 t3 with l = u + 1/2: -t3 u^2 + (a - t3) u + (a/2 - t3/4)
          const double ax = a.x, ay = a.y;
          GridEnt ge;
          ge.ux = (float)ox - 0.5f;
          ge.uy = (float)oy - 0.5f;
          ge.sg = mir ? -1.f : 1.f;
          ge.a0 = (float)(kL2E * (0.5 * ax - 0.25 * t3x + 0.5 * ay - 0.25 * t3y));
          ge.a1x = (float)(kL2E * (ax - t3x));
          ge.a2x = (float)(-kL2E * t3x);
          ge.a1y = (float)(kL2E * (ay - t3y));
          ge.a2y = (float)(-kL2E * t3y);
          ge.dx = ew * r.x;
          ge.dy = mir ? ew * r.y : -ew * r.y;
          ge.pad0 = 0.f;
          ge.pad1 = 0.f;
          er[q] = ge;
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
        const GridEnt ge = er[q];
#pragma unroll
        for (int c = 0; c < CPT; c++) {
          const float u = fmaf(ge.sg, fcc[c], ge.ux), up = fmaf(ge.sg, frr[c], ge.uy);
          const float ex = fmaf(fmaf(ge.a2x, u, ge.a1x), u, ge.a0);
          const float e = fmaf(fmaf(ge.a2y, up, ge.a1y), up, ex);
          const float wgt = (fabsf(u) < fmsp && fabsf(up) < fmsp) ? __builtin_amdgcn_exp2f(e) : 0.f;
          acc[c].x = fmaf(wgt, ge.dx, acc[c].x);
          acc[c].y = fmaf(wgt, ge.dy, acc[c].y);
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

// The control step's decision for one trial round, taken by thread 0 of every
// workgroup with the same arithmetic everywhere.  sum: the round's reduced
// partials, trial k at sum[k kPTrial + slot] (slots kPChi, kPLip, ...  in the
// order chi^2(xnew), |W A d|^2, |d|^2, |xnew|_1, reg(xnew), TSV(d), d_X(xnew),
// d_X(d)); chi_add is added to every trial's chi^2 (the constant |W V|^2 of
// the Toeplitz loop, 0 otherwise).  The trials of a round are taken in order,
// exactly as the reference's backtracking loop (c/mfista_nufft_lib.c:692-723)
// takes them: the first that passes is accepted; one that fails multiplies c
// by eta.
__device__ __forceinline__ void round_decide(const PArgs& A, BlockState& bs, double* ring, const double* sum, bool pr,
                                    bool tv, bool tsv, bool movie, int lb, int prologue, double chi_add) {
  const int it = bs.iter;
  if (prologue) {
    ring[it % (kTD + 1)] = bs.costtmp;
    if (lb == 0) GP(A.cost)[it] = bs.costtmp;
    bs.prologue = 0;
  }
  bs.rounds += 1;
  const double etaf = (it == 0) ? A.eta_first : kEta;
  const long cap = (long)A.maxiter * 64 + 100000;
  const int K = pr ? 2 : 1;
  for (int k = 0; k < K; k++) {
    const double* sk = sum + k * kPTrial;
    const double chi_x = sk[0] + chi_add, lip = sk[1], s_d2 = sk[2], l1n = sk[3], regn = sk[4], tsvd = sk[5];
    double F = chi_x / 2 + ((!tv && tsv) ? A.lambda_tsv * regn : 0.0);
    double lhs = lip / 2 + ((!tv && tsv) ? A.lambda_tsv * tsvd : 0.0);
    if (movie) {  // + lambda_t d_X(xnew); d_X is quadratic, so its step term is d_X(d)
      F += A.lambda_t * sk[6];
      lhs += A.lambda_t * sk[7];
    }
    const double rhs = bs.c * s_d2 / 2;
    bs.trial += 1;
    bs.total_trials += 1;
    if (bs.total_trials > cap) {  // cannot happen for a convergent run; bounds the launch
      bs.status = 2;
      bs.stop = 1;
      bs.ITER = it + 1;
      bs.mom_pending = 0;
      break;
    }
    if (!isfinite(F) || !isfinite(lhs)) {
      bs.status = 1;
      bs.stop = 1;
      bs.ITER = it + 1;
      bs.mom_pending = 0;
      break;
    }
    const bool ok = lhs <= rhs;
    if (!ok && bs.trial < A.maxiter) {
      bs.c *= etaf;
      continue;  // the next trial of this round, or a new round
    }
    if (!ok) bs.c *= etaf;
    if (bs.trial > bs.max_trials) bs.max_trials = bs.trial;
    bs.kacc = k;
    bs.c /= kEta;
    const double mu = bs.mu;
    const double munew = (1 + sqrt(1 + 4 * mu * mu)) / 2;
    F += A.lambda_l1 * l1n;
    if (tv) F += A.lambda_tv * regn;
    bool stop = false;
    const double cit = ring[it % (kTD + 1)];
    if (F < cit) {
      bs.costtmp = F;
      bs.mom_a = (1 - mu) / munew;
      bs.mom_b = 1 + ((mu - 1) / munew);
      bs.accept = 1;
      bs.l1_x = l1n;
    } else {
      bs.mom_a = 1 - (mu / munew);
      bs.mom_b = mu / munew;
      bs.accept = 0;
      // c/mfista_nufft_lib.c:737 (the movie solver has no such rule)
      if (!movie && it > 1 && bs.l1_x == 0) stop = true;
    }
    if (!stop && it >= kMinIter) {
      const double d = ring[(it - kTD) % (kTD + 1)] - cit;
      stop = (A.rel_tol > 0) ? (d <= A.rel_tol * cit) : (d < A.eps_abs);
    }
    bs.mu = munew;
    bs.trial = 0;
    bs.mom_pending = 1;
    if (stop) {
      bs.stop = 1;
      bs.ITER = it + 1;
    } else {
      bs.iter = it + 1;
      if (bs.iter >= A.maxiter) {
        bs.stop = 1;
        bs.ITER = A.maxiter;
      } else {
        bs.prologue = 1;
      }
    }
    break;
  }
}

// LDS bytes of the batched gridder's entry records
__host__ __device__ inline size_t grid_lds_bytes(int TX, int TY) {
  (void)TX;
  (void)TY;
  return (size_t)MF_BLOCK * sizeof(GridEnt) + (size_t)kGB * 16;
}

// Visibility records of the paired degrid tasks cached in LDS per workgroup
// (64 bytes each): the first 3 tasks of 20 visibilities (12 taps) or 8 tasks
// of 8 (24 taps) of each workgroup; later tasks load their records.
constexpr int kDVisCap = 64;

// LDS bytes of the staged twiddle (FP32 and FP64, both axes) and E4 tables
__host__ __device__ inline size_t table_lds_bytes(int Lx, int Ly, int NX, int NY) {
  return (size_t)(Lx + Ly) * (16 + 8) + (size_t)(NX + NY) * (8 + 4);
}

// The solve of one problem by workgroups lb = 0 .. nb-1 of a cooperative
// launch (all of the launch for a single solve, a share of it in a batch).
// TAB: the twiddle and E4 tables are staged in LDS.  A compile-time choice, so
// that every table read is an LDS or a GLOBAL access; a pointer that may hold
// either compiles to FLAT accesses (see gld in common.h).
template <int MSPC, bool TAB, int CPT>
__device__ __forceinline__ void mfista_body(const PArgs& A, const int lb, const unsigned nb) {
  extern __shared__ double4 smem_d4[];
  __shared__ double red[16 * kPW];
  __shared__ BlockState bs;
  __shared__ double ring[kTD + 1];
  __shared__ BarCtx B;
  const GridGeom& g = A.g;
  const int NX = g.NX, NY = g.NY, n = NX * NY, W = NX, H = NY;
  const bool tsv = A.lambda_tsv > 0.0;
  const bool tv = A.tv_mode != 0;
  unsigned long long epoch = 0;  // barriers passed by this workgroup
  const bool movie = A.ngb > 0;
  const unsigned ngrp = movie ? (unsigned)A.ngb : nb;  // workgroups of the barrier and the reductions
  // this workgroup's column of the slot-major partial table
  double* prow = GP(A.part) + (movie ? A.gb0 + (int)lb : (int)lb);
  const size_t pstr = ngrp;

  if (!bar_setup(B, GP(A.bar), ngrp, A.hier_barrier)) return;
  if (threadIdx.x == 0) {
    B.trace = A.trace;
    B.ttag = A.ttag;
    B.tcap = A.tcap;
    B.trow = movie ? A.gb0 + (int)lb : (int)lb;
    B.tnrow = (int)ngrp;
  }
  __syncthreads();

  // LDS: [tables][stage scratch]
  char* lds = reinterpret_cast<char*>(smem_d4);
  const double2 *twxd = GP(A.twxd), *twyd = GP(A.twyd);
  const float2 *twxf = GP(A.twxf), *twyf = GP(A.twyf);
  const double *e4xd = GP(A.e4xd), *e4yd = GP(A.e4yd);
  const float *e4xf = GP(A.e4xf), *e4yf = GP(A.e4yf);
  if constexpr (TAB) {
    double2* t64x = reinterpret_cast<double2*>(lds);
    double2* t64y = t64x + g.Lx;
    float2* t32x = reinterpret_cast<float2*>(t64y + g.Ly);
    float2* t32y = t32x + g.Lx;
    double* q64x = reinterpret_cast<double*>(t32y + g.Ly);
    double* q64y = q64x + NX;
    float* q32x = reinterpret_cast<float*>(q64y + NY);
    float* q32y = q32x + NX;
    for (int i = threadIdx.x; i < g.Lx; i += blockDim.x) {
      t64x[i] = GP(A.twxd)[i];
      t32x[i] = GP(A.twxf)[i];
    }
    for (int i = threadIdx.x; i < g.Ly; i += blockDim.x) {
      t64y[i] = GP(A.twyd)[i];
      t32y[i] = GP(A.twyf)[i];
    }
    for (int i = threadIdx.x; i < NX; i += blockDim.x) {
      q64x[i] = GP(A.e4xd)[i];
      q32x[i] = GP(A.e4xf)[i];
    }
    for (int i = threadIdx.x; i < NY; i += blockDim.x) {
      q64y[i] = GP(A.e4yd)[i];
      q32y[i] = GP(A.e4yf)[i];
    }
    twxd = t64x;
    twyd = t64y;
    twxf = t32x;
    twyf = t32y;
    e4xd = q64x;
    e4yd = q64y;
    e4xf = q32x;
    e4yf = q32y;
    lds += (table_lds_bytes(g.Lx, g.Ly, NX, NY) + 255) & ~(size_t)255;
    __syncthreads();
  }
  // the visibility records of this workgroup's degrid tasks
  constexpr int kVPT = PairShape<MSPC>::VPT;  // visibilities per paired degrid task
  constexpr int kTC = kDVisCap / kVPT;         // tasks cached
  const int n_degp = (g.M + kVPT - 1) / kVPT;  // paired degrid tasks
  DVis* dvc = reinterpret_cast<DVis*>(lds);
  lds += kDVisCap * sizeof(DVis);
  for (int q = threadIdx.x; q < kTC * kVPT; q += blockDim.x) {
    const int k = q / kVPT, t = lb + k * (int)nb;
    if (t < n_degp) dvc[q] = dvis_load(A.pd, g, t * kVPT + q % kVPT);
  }
  __syncthreads();
  float2* sm32 = reinterpret_cast<float2*>(lds);
  double2* sm64 = reinterpret_cast<double2*>(lds);

  // ---------------- initial cost and models: W A x0 -> vm[0] (x), vm[2] (z)
  {
    double st[4] = {0, 0, 0, 0};
    double sdx[1] = {0};
    for (int t = lb; t < A.n_row; t += nb) {
      auto load = [&](int i, int j) -> double {
        const int f = i * NY + j;
        const double xc = GP(A.x)[f];
        GP(A.z)[f] = xc;
        st[0] += fabs(xc);
        if (GP(A.x_next)) {
          const double e = xc - GP(A.x_next)[f];
          sdx[0] += A.wt_next * e * e;
        }
        const int c = f % W, r = f / W;
        double dh = 0, dv = 0;
        if (c < W - 1) dh = xc - GP(A.x)[f + 1];
        if (r < H - 1) dv = xc - GP(A.x)[f + W];
        st[1] += dh * dh + dv * dv;
        if (c < W - 1 && r < H - 1)
          st[2] += sqrt(dh * dh + dv * dv);
        else if (c < W - 1)
          st[2] += fabs(dh);
        else if (r < H - 1)
          st[2] += fabs(dv);
        return xc * e4xd[i] * e4yd[j];
      };
      st_rowfft_fwd<double>(t, 1, NX, NY, g.Ly, NY / 2, A.Py, twyd, GP(A.T64), sm64, load);
    }
    block_sum<4>(st, red);
    block_sum<1>(sdx, red);
    if (threadIdx.x == 0) {
      prow[(kPIl1) * pstr] = st[0];
      prow[(kPItsv) * pstr] = st[1];
      prow[(kPItv) * pstr] = st[2];
      prow[(kPIdx) * pstr] = sdx[0];
    }
  }
  if (!grid_sync(B, epoch, 1)) return;
  for (int t = lb; t < A.n_ky; t += nb)
    st_colfft_fwd<double>(t, 1, GP(A.T64), GP(A.Gf64), NX, NY + 1, g.Lx, NX / 2, 0, g.MSP, A.Px, twxd, sm64);
  if (!grid_sync(B, epoch, 2)) return;
  {
    double v[2] = {0, 0};
    for (int t = lb; t < A.n_deg; t += nb) st_degrid_model<MSPC>(t, g, A.pd, GP(A.Gf64), nullptr, GP(A.vm[0]), GP(A.vm[2]), v);
    block_sum<2>(v, red);
    if (threadIdx.x == 0) prow[(kPIchi) * pstr] = v[0];
  }
  if (!grid_sync(B, epoch, 3)) return;
  {
    const int slot[5] = {kPIchi, kPIl1, kPItsv, kPItv, kPIdx};
    double sum[5];
    part_reduce<5>(GP(A.part), ngrp, slot, sum, red);
    if (threadIdx.x == 0) {
      double c0 = sum[0] / 2 + A.lambda_l1 * sum[1];
      if (tv)
        c0 += A.lambda_tv * sum[3];
      else if (tsv)
        c0 += A.lambda_tsv * sum[2];
      if (movie) c0 += A.lambda_t * sum[4];
      bs.c = A.cinit;
      bs.mu = 1.0;
      bs.costtmp = c0;
      bs.l1_x = sum[1];
      bs.iter = 0;
      bs.trial = 0;
      bs.accept = 0;
      bs.mom_pending = 0;
      bs.stop = 0;
      bs.ITER = 0;
      bs.max_trials = 0;
      bs.total_trials = 0;
      bs.status = 0;
      bs.prologue = 1;
      bs.ix = 0;
      bs.in = 1;
      bs.iz = 2;
      bs.in1 = 3;
      bs.kacc = 0;
      bs.rounds = 0;
    }
    __syncthreads();
  }

  // ---------------- the MFISTA loop
  while (true) {
    const int prologue = bs.prologue;
    if (prologue) {
      // P1: momentum (images and models) and gridding of the residual at z
      const bool mom = bs.mom_pending != 0;
      const double ma = bs.mom_a, mb = bs.mom_b;
      const bool acc = bs.accept != 0;
      const int ka = bs.kacc;  // the accepted trial's xnew and model
      const double* xnw = GP(A.xn[ka]);
      const double2* mx = GP(A.vm[bs.ix]);
      const double2* mn = GP(A.vm[ka ? bs.in1 : bs.in]);
      double2* mz = GP(A.vm[bs.iz]);
      if (mom) {
        for (int f = lb * blockDim.x + threadIdx.x; f < n; f += nb * blockDim.x) {
          const double xf = GP(A.x)[f], xn = xnw[f];
          GP(A.z)[f] = ma * xf + mb * xn;
          if (acc) GP(A.x)[f] = xn;
        }
        for (int s = lb * blockDim.x + threadIdx.x; s < g.M; s += nb * blockDim.x) {
          const double2 a = mx[s], b = mn[s];
          mz[s] = make_double2(ma * a.x + mb * b.x, ma * a.y + mb * b.y);
        }
      }
      auto resid = [&](int s) -> float2 {
        const double w = GP(A.pd.w)[s];
        const double2 V = GP(A.pd.vis)[s];
        double2 m;
        if (mom) {
          const double2 a = mx[s], b = mn[s];
          m = make_double2(ma * a.x + mb * b.x, ma * a.y + mb * b.y);
        } else {
          m = mz[s];
        }
        return make_float2((float)(V.x * w - m.x), (float)(V.y * w - m.y));
      };
      grid_chunks<CPT>(g, A.pf, resid, GP(A.chunks), GP(A.granges)[lb],
                                              GP(A.granges)[lb + 1], GP(A.ent),
                                              GP(A.Hg), lds);
      if (!grid_sync(B, epoch, 4)) return;
      if (threadIdx.x == 0 && mom && acc) {  // W A x <- W A xnew: swap roles
        int& rn = ka ? bs.in1 : bs.in;
        const int tmp = bs.ix;
        bs.ix = rn;
        rn = tmp;
      }
      // P2: inverse column pass (clears Hg behind it)
      for (int t = lb; t < A.n_ky; t += nb)
        st_colfft_inv<float>(t, 1, GP(A.Hg), GP(A.U), NX, NY + 1, g.Lx, NX / 2, 0, A.Px, twxf, sm32);
      if (!grid_sync(B, epoch, 5)) return;
    }
    const int trial_first = prologue;
    const double c = bs.c;
    const float cinv = (float)(1.0 / c);
    const float eta = (float)(A.lambda_l1 / c);
    const float lam = (float)(A.lambda_tv / c);
    const float lt = (float)A.lambda_tsv, sc = (float)A.adj_scale;
    // gradient store of the c2r pass: g = Re A^H W r at z (scaled, E4 crop)
    // minus lambda_tsv grad TSV(z)
    auto gstore = [&](int i, int j, float v) {
      const int f = i * NY + j;
      float gv = v * sc * e4xf[i] * e4yf[j];
      if (tsv) {
        const int cc = f % W, r = f / W;
        const double zc = GP(A.z)[f];
        double d = 0.0;
        if (cc < W - 1) d += zc - GP(A.z)[f + 1];
        if (cc > 0) d += zc - GP(A.z)[f - 1];
        if (r < H - 1) d += zc - GP(A.z)[f + W];
        if (r > 0) d += zc - GP(A.z)[f - W];
        gv = gv - lt * (float)(2.0 * d);
      }
      if (movie) {  // - lambda_t dt_X(z) (matlab/lib/dt_X.m)
        const double zc = GP(A.z)[f];
        double d = 0.0;
        if (GP(A.z_next)) d += A.wt_next * (zc - GP(A.z_next)[f]);
        if (GP(A.z_prev)) d += A.wt_prev * (zc - GP(A.z_prev)[f]);
        gv = gv - (float)(A.lambda_t * 2.0 * d);
      }
      GP(A.gr)[f] = gv;
    };
    if (tv) {
      if (trial_first) {
        for (int t = lb; t < A.n_row; t += nb)
          st_rowfft_c2r<float>(t, 1, GP(A.U), NX, NY, g.Ly, NY / 2, A.Py, twyf, sm32, gstore);
        if (!grid_sync(B, epoch, 7)) return;
      }
      // FGP iterations (c/mfista_TV_lib.c:189-384), b = z + g/c (nonneg: -l1/c + g/c + z)
      const float bshift = (A.nonneg == 1) ? (float)(-A.lambda_l1 / c) : 0.f;
      const float a3 = 1.f / (8.f * lam);
      float *ri = GP(A.rA), *si = GP(A.sA), *ro = GP(A.rB), *so = GP(A.sB);
      for (int k = 0; k < A.fgp_iter; k++) {
        const float c1 = (float)GP(A.fgp_c)[2 * k], c2 = (float)GP(A.fgp_c)[2 * k + 1];
        const int first = (k == 0);
        // one FGP update of pixel (i, j): x at (i, j), (i + 1, j), (i, j + 1)
        // from b and the previous duals, then the new duals
        auto fgp_pixel = [&](const int i, const int j) {
          auto xat = [&](int ii, int jj) -> float {
            const int ff = jj * W + ii;
            const float bb = bshift + cinv * GP(A.gr)[ff] + (float)GP(A.z)[ff];
            const float L = first ? 0.f : lpq_f(ri, si, ii, jj, W, H);
            return fgp_proj(bb - lam * L, eta, A.nonneg == 1, GP(A.box) && GP(A.box)[ff] == 0.0f);
          };
          const float xc = xat(i, j);
          float rr = 0.f, ss = 0.f;
          const bool hp = i < W - 1, hq = j < H - 1;
          if (hp) rr = (first ? 0.f : ri[(W - 1) * j + i]) + (xc - xat(i + 1, j)) * a3;
          if (hq) ss = (first ? 0.f : si[W * j + i]) + (xc - xat(i, j + 1)) * a3;
          float np_ = 0.f, nq_ = 0.f;
          if (hp && hq) {
            float tt = sqrtf(rr * rr + ss * ss);
            if (tt < 1.f) tt = 1.f;
            np_ = rr / tt;
            nq_ = ss / tt;
          } else if (hp) {
            float tt = fabsf(rr);
            if (tt < 1.f) tt = 1.f;
            np_ = rr / tt;
          } else if (hq) {
            float tt = fabsf(ss);
            if (tt < 1.f) tt = 1.f;
            nq_ = ss / tt;
          }
          if (hp) {
            const int kk = (W - 1) * j + i;
            const float pold = first ? 0.f : GP(A.pp)[kk];
            ro[kk] = np_ * c1 + c2 * pold;
            GP(A.pp)[kk] = np_;
          }
          if (hq) {
            const int kk = W * j + i;
            const float qold = first ? 0.f : GP(A.qq)[kk];
            so[kk] = nq_ * c1 + c2 * qold;
            GP(A.qq)[kk] = nq_;
          }
        };
        // pixel -> thread mapping (the FGP block shape, PArgs::fgp_tw): rows
        // of the flat image one pixel per thread (0), or tiles of fgp_tw x
        // (256 / fgp_tw) pixels dealt to the workgroups; every pixel's update
        // is the same either way
        if (A.fgp_tw == 0) {
          for (int f = lb * blockDim.x + threadIdx.x; f < n; f += nb * blockDim.x) fgp_pixel(f % W, f / W);
        } else {
          const int tw = A.fgp_tw, th = MF_BLOCK / tw;
          const int ntx = (W + tw - 1) / tw, nty = (H + th - 1) / th;
          const int ti0 = threadIdx.x % tw, tj0 = threadIdx.x / tw;
          for (int t = lb; t < ntx * nty; t += nb) {
            const int i = (t % ntx) * tw + ti0, j = (t / ntx) * th + tj0;
            if (i < W && j < H) fgp_pixel(i, j);
          }
        }
        if (!grid_sync(B, epoch, 6)) return;
        float* tmp = ri;
        ri = ro;
        ro = tmp;
        tmp = si;
        si = so;
        so = tmp;
      }
    }
    // T1: [c2r + gradient on the first trial] -> xnew (prox or final FGP step),
    // d = xnew - z (FP64), partials, FP64 forward row pass of d.  A paired
    // round (L1+TSV) computes the steps of c and c * eta side by side.
    const bool pr = A.pair && !tv;
    if (pr) {
      const double etaf = (bs.iter == 0) ? A.eta_first : kEta;
      const double c1 = c * etaf;  // the value the control step gives c after a failed trial
      const double ci0 = 1.0 / c, e0 = A.lambda_l1 / c, ci1 = 1.0 / c1, e1 = A.lambda_l1 / c1;
      double v[4] = {0, 0, 0, 0};
      // The first trial of an iteration (outside a movie, images up to 256^2)
      // loads the c2r's input, the z values of its TSV gradient and the box
      // together and stages them in LDS; the prox then reads the gradient and
      // z from LDS instead of from global memory.  The staging area is the
      // upper half of the FP64 row-pass scratch, unused until that pass's
      // first FFT stage.  Measured at 256^2: T1 15.8 -> 14.9 us per round.
      constexpr int kZ = 4, kB = 2;  // staged z and box values per thread
      const int zspan = 2 * NY + 2 * W;
      const bool fused = trial_first && !movie && zspan <= kZ * MF_BLOCK && 2 * NY <= kB * MF_BLOCK &&
                         (size_t)zspan * 8 + (size_t)4 * NY * 4 <= (size_t)2 * g.Ly * sizeof(double2);
      double* zs = reinterpret_cast<double*>(sm64 + 2 * g.Ly);
      float* gs = reinterpret_cast<float*>(zs + zspan);
      float* bxs = gs + 2 * NY;
      const float* boxp = GP(A.box);
      for (int t = lb; t < A.n_row; t += nb) {
        if (fused) {
          const int f_lo = 2 * t * NY, f_hi = min(n, f_lo + 2 * NY);
          const int zb = max(0, f_lo - W), ze = min(n, f_hi + W);
          double zr[kZ];
          float br[kB];
          auto issue = [&]() {
#pragma unroll
            for (int k = 0; k < kZ; k++) {
              const int f = zb + (int)threadIdx.x + k * MF_BLOCK;
              zr[k] = (f < ze) ? gld(GP(A.z) + f) : 0.0;
            }
#pragma unroll
            for (int k = 0; k < kB; k++) {
              const int f = f_lo + (int)threadIdx.x + k * MF_BLOCK;
              br[k] = (boxp && f < f_hi) ? gld(boxp + f) : 1.0f;
            }
          };
          auto commit = [&]() {
#pragma unroll
            for (int k = 0; k < kZ; k++) {
              const int f = zb + (int)threadIdx.x + k * MF_BLOCK;
              if (f < ze) zs[f - zb] = zr[k];
            }
#pragma unroll
            for (int k = 0; k < kB; k++) {
              const int q = (int)threadIdx.x + k * MF_BLOCK;
              if (q < 2 * NY) bxs[q] = br[k];
            }
          };
          auto gstore_s = [&](int i, int j, float vv) {
            const int f = i * NY + j;
            float gv = vv * sc * e4xf[i] * e4yf[j];
            if (tsv) {
              const int cc = f % W, r = f / W;
              const double zc = zs[f - zb];
              double d = 0.0;
              if (cc < W - 1) d += zc - zs[f + 1 - zb];
              if (cc > 0) d += zc - zs[f - 1 - zb];
              if (r < H - 1) d += zc - zs[f + W - zb];
              if (r > 0) d += zc - zs[f - W - zb];
              gv = gv - lt * (float)(2.0 * d);
            }
            GP(A.gr)[f] = gv;
            gs[f - f_lo] = gv;
          };
          st_rowfft_c2r_pf<float>(t, GP(A.U), NX, NY, g.Ly, NY / 2, A.Py, twyf, sm32, gstore_s, issue, commit);
          auto load2s = [&](int i, int j, double& o0, double& o1) {
            const int f = i * NY + j;
            const double zf = zs[f - zb], gf = (double)gs[f - f_lo];
            const bool boxed = boxp && bxs[f - f_lo] == 0.0f;
            const double xa = prox_p<double>(gf * ci0 + zf, e0, A.nonneg, boxed);
            const double xb = prox_p<double>(gf * ci1 + zf, e1, A.nonneg, boxed);
            GP(A.xn[0])[f] = xa;
            GP(A.xn[1])[f] = xb;
            const double da = xa - zf, db = xb - zf;
            v[0] += da * da;
            v[1] += fabs(xa);
            v[2] += db * db;
            v[3] += fabs(xb);
            o0 = da * e4xd[i] * e4yd[j];
            o1 = db * e4xd[i] * e4yd[j];
          };
          st_rowfft_fwd_pair<double>(t, NX, NY, g.Ly, NY / 2, A.Py, twyd, GP(A.T64), GP(A.T64b), sm64, load2s);
          continue;
        }
        if (trial_first) st_rowfft_c2r<float>(t, 1, GP(A.U), NX, NY, g.Ly, NY / 2, A.Py, twyf, sm32, gstore);
        auto load2 = [&](int i, int j, double& o0, double& o1) {
          const int f = i * NY + j;
          const double zf = GP(A.z)[f], gf = (double)GP(A.gr)[f];
          const bool boxed = GP(A.box) && GP(A.box)[f] == 0.0f;
          const double xa = prox_p<double>(gf * ci0 + zf, e0, A.nonneg, boxed);
          const double xb = prox_p<double>(gf * ci1 + zf, e1, A.nonneg, boxed);
          GP(A.xn[0])[f] = xa;
          GP(A.xn[1])[f] = xb;
          const double da = xa - zf, db = xb - zf;
          v[0] += da * da;
          v[1] += fabs(xa);
          v[2] += db * db;
          v[3] += fabs(xb);
          o0 = da * e4xd[i] * e4yd[j];
          o1 = db * e4xd[i] * e4yd[j];
        };
        st_rowfft_fwd_pair<double>(t, NX, NY, g.Ly, NY / 2, A.Py, twyd, GP(A.T64), GP(A.T64b), sm64, load2);
      }
      block_sum<4>(v, red);
      if (threadIdx.x == 0) {
        prow[(kPD2) * pstr] = v[0];
        prow[(kPL1) * pstr] = v[1];
        prow[(kPTrial + kPD2) * pstr] = v[2];
        prow[(kPTrial + kPL1) * pstr] = v[3];
      }
    } else {
      double v[2] = {0, 0};
      const double cinv_d = 1.0 / c, eta_d = A.lambda_l1 / c, lam_d = A.lambda_tv / c;
      const double bshift = (tv && A.nonneg == 1) ? -eta_d : 0.0;
      const bool have_dual = tv && A.fgp_iter > 0;
      for (int t = lb; t < A.n_row; t += nb) {
        if (!tv && trial_first)
          st_rowfft_c2r<float>(t, 1, GP(A.U), NX, NY, g.Ly, NY / 2, A.Py, twyf, sm32, gstore);
        auto load = [&](int i, int j) -> double {
          const int f = i * NY + j;
          const double zf = GP(A.z)[f];
          const bool boxed = GP(A.box) && GP(A.box)[f] == 0.0f;
          double xn;
          if (!tv) {
            xn = prox_p<double>((double)GP(A.gr)[f] * cinv_d + zf, eta_d, A.nonneg, boxed);
          } else {
            const int ii = f % W, jj = f / W;
            const double bb = bshift + cinv_d * (double)GP(A.gr)[f] + zf;
            const double L = have_dual ? (double)lpq_f(GP(A.pp), GP(A.qq), ii, jj, W, H) : 0.0;
            xn = fgp_proj_d(bb - lam_d * L, eta_d, A.nonneg == 1, boxed);
          }
          GP(A.xn[0])[f] = xn;
          const double d = xn - zf;
          v[0] += d * d;
          v[1] += fabs(xn);
          return d * e4xd[i] * e4yd[j];
        };
        st_rowfft_fwd<double>(t, 1, NX, NY, g.Ly, NY / 2, A.Py, twyd, GP(A.T64), sm64, load);
      }
      block_sum<2>(v, red);
      if (threadIdx.x == 0) {
        prow[(kPD2) * pstr] = v[0];
        prow[(kPL1) * pstr] = v[1];
      }
    }
    if (!grid_sync(B, epoch, 8)) return;
    // T2: FP64 column pass of d; regulariser partials of xnew and d.  When at
    // least half as many workgroups as pixels / 256 have no column task, those
    // workgroups take the partials and the column-task workgroups skip them:
    // one dependent round trip fewer on the stage's critical path.
    const int nreg = (int)nb - A.n_ky;
    const bool split = nreg > 0 && 2 * (long)nreg * blockDim.x >= n;
    const bool do_reg = !split || (int)lb >= A.n_ky;
    const int rg0 = split ? (int)lb - A.n_ky : (int)lb;
    const int rgn = split ? nreg : (int)nb;
    if (pr) {
      for (int t = lb; t < A.n_ky; t += nb)
        st_colfft_fwd_pair<double>(t, GP(A.T64), GP(A.T64b), GP(A.Gf64), GP(A.Gf64b), NX, NY + 1, g.Lx, NX / 2, g.MSP, A.Px,
                                   twxd, sm64);
      double r[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      for (int f = rg0 * blockDim.x + threadIdx.x; do_reg && f < n; f += rgn * blockDim.x) {
        reg_terms(GP(A.xn[0]), GP(A.z), f, W, H, false, r);
        reg_terms(GP(A.xn[1]), GP(A.z), f, W, H, false, r + 2);
        if (GP(A.x_next)) {
#pragma unroll
          for (int k = 0; k < 2; k++) temporal_terms(A, k, f, r + 4 + 2 * k);
        }
      }
      block_sum<8>(r, red);
      if (threadIdx.x == 0) {
        prow[(kPReg) * pstr] = r[0];
        prow[(kPTsvd) * pstr] = r[1];
        prow[(kPTrial + kPReg) * pstr] = r[2];
        prow[(kPTrial + kPTsvd) * pstr] = r[3];
        prow[(kPTx) * pstr] = r[4];
        prow[(kPTd) * pstr] = r[5];
        prow[(kPTrial + kPTx) * pstr] = r[6];
        prow[(kPTrial + kPTd) * pstr] = r[7];
      }
    } else {
      for (int t = lb; t < A.n_ky; t += nb)
        st_colfft_fwd<double>(t, 1, GP(A.T64), GP(A.Gf64), NX, NY + 1, g.Lx, NX / 2, 0, g.MSP, A.Px, twxd, sm64);
      double r[4] = {0, 0, 0, 0};
      for (int f = rg0 * blockDim.x + threadIdx.x; do_reg && f < n; f += rgn * blockDim.x) {
        reg_terms(GP(A.xn[0]), GP(A.z), f, W, H, tv, r);
        if (GP(A.x_next)) temporal_terms(A, 0, f, r + 2);
      }
      block_sum<4>(r, red);
      if (threadIdx.x == 0) {
        prow[(kPReg) * pstr] = r[0];
        prow[(kPTsvd) * pstr] = r[1];
        prow[(kPTx) * pstr] = r[2];
        prow[(kPTd) * pstr] = r[3];
      }
    }
    if (!grid_sync(B, epoch, 9)) return;
    // T3: FP64 degrid of d: W A xnew = W A z + W A d, chi^2(xnew), |W A d|^2
    {
      const double2* mz = GP(A.vm[bs.iz]);
      if (pr) {
        double v[4] = {0, 0, 0, 0};
        for (int t = lb, k = 0; t < n_degp; t += nb, k++)
          st_degrid_model_pair<MSPC>(t, g, A.pd, k < kTC ? dvc + k * kVPT : nullptr, GP(A.Gf64), GP(A.Gf64b), mz,
                                     GP(A.vm[bs.in]), GP(A.vm[bs.in1]), v);
        block_sum<4>(v, red);
        if (threadIdx.x == 0) {
          prow[(kPChi) * pstr] = v[0];
          prow[(kPLip) * pstr] = v[1];
          prow[(kPTrial + kPChi) * pstr] = v[2];
          prow[(kPTrial + kPLip) * pstr] = v[3];
        }
      } else {
        double v[2] = {0, 0};
        for (int t = lb; t < A.n_deg; t += nb)
          st_degrid_model<MSPC>(t, g, A.pd, GP(A.Gf64), mz, GP(A.vm[bs.in]), nullptr, v);
        block_sum<2>(v, red);
        if (threadIdx.x == 0) {
          prow[(kPChi) * pstr] = v[0];
          prow[(kPLip) * pstr] = v[1];
        }
      }
    }
    if (!grid_sync(B, epoch, 10)) return;

    // ---------------- control (every workgroup, identical arithmetic): the
    // round's partials reduced in a fixed order, then round_decide
    double sum[2 * kPTrial];
    if (pr && (movie || MSPC > 8)) {
      const int sl[2 * kPTrial] = {kPChi, kPLip, kPD2, kPL1, kPReg, kPTsvd, kPTx, kPTd,
                                   kPTrial + kPChi, kPTrial + kPLip, kPTrial + kPD2, kPTrial + kPL1,
                                   kPTrial + kPReg, kPTrial + kPTsvd, kPTrial + kPTx, kPTrial + kPTd};
      part_reduce<2 * kPTrial>(GP(A.part), ngrp, sl, sum, red);
    } else if (pr && MSPC <= 8) {  // outside a movie the temporal slots (kPTx, kPTd) are zero: not read
      // (12 taps only: the 24-tap kernel spills more with a second reduction shape)
      const int sl[12] = {kPChi, kPLip, kPD2, kPL1, kPReg, kPTsvd,
                          kPTrial + kPChi, kPTrial + kPLip, kPTrial + kPD2, kPTrial + kPL1,
                          kPTrial + kPReg, kPTrial + kPTsvd};
      double s1[12];
      part_reduce<12>(GP(A.part), ngrp, sl, s1, red);
#pragma unroll
      for (int k = 0; k < 2; k++) {
#pragma unroll
        for (int q = 0; q < 6; q++) sum[k * kPTrial + q] = s1[k * 6 + q];
        sum[k * kPTrial + kPTx] = 0.0;
        sum[k * kPTrial + kPTd] = 0.0;
      }
    } else {
      const int sl[kPTrial] = {kPChi, kPLip, kPD2, kPL1, kPReg, kPTsvd, kPTx, kPTd};
      double s1[kPTrial];
      part_reduce<kPTrial>(GP(A.part), ngrp, sl, s1, red);
#pragma unroll
      for (int q = 0; q < kPTrial; q++) sum[q] = s1[q];
    }
    if (threadIdx.x == 0) round_decide(A, bs, ring, sum, pr, tv, tsv, movie, (int)lb, prologue, 0.0);
    __syncthreads();
    if (bs.stop) break;
  }

  // the accepted image of the last iteration (the reference's x = xnew)
  if (bs.mom_pending && bs.accept) {
    const double* xnw = GP(A.xn[bs.kacc]);
    for (int f = lb * blockDim.x + threadIdx.x; f < n; f += nb * blockDim.x) GP(A.x)[f] = xnw[f];
  }
  if (lb == 0 && threadIdx.x == 0) {
    DevState* S = A.S;
    S->c = bs.c;
    S->mu = bs.mu;
    S->costtmp = bs.costtmp;
    S->ITER = bs.ITER;
    S->iter = bs.iter;
    S->max_trials = bs.max_trials;
    S->total_trials = bs.total_trials;
    S->status = bs.status;
    S->rounds = bs.rounds;
    S->done = 1;
  }
}

// The arguments are read from device memory rather than the kernel-argument
// segment: passed by value they would be held in scalar registers, whose
// spills into vector registers cost the 24-tap kernel its second workgroup
// per CU.  The 24-tap kernel also needs an explicit bound of two waves per
// SIMD (two workgroups per CU) to keep them; the 12-tap kernel fits without one,
// and the bound's scratch spills in its control step cost more than they save.
template <int MSPC>
constexpr int kMinWaves = (MSPC > 8) ? 2 : 1;

template <int MSPC, bool TAB, int CPT>
__global__ void __launch_bounds__(MF_BLOCK, kMinWaves<MSPC>) k_mfista_persistent(const PArgs* __restrict__ A) {
  mfista_body<MSPC, TAB, CPT>(*A, blockIdx.x, gridDim.x);
}

// A batch of independent solves in one cooperative launch: blk[b] =
// (problem, workgroup index within the problem) for workgroup b.
template <int MSPC, bool TAB, int CPT>
__global__ void __launch_bounds__(MF_BLOCK, kMinWaves<MSPC>) k_mfista_batch(const PArgs* __restrict__ args,
                                                                             const int2* __restrict__ blk) {
  const int2 pb = blk[blockIdx.x];
  if (pb.x < 0) return;  // a workgroup slot no problem was dealt
  const PArgs& A = args[pb.x];
  mfista_body<MSPC, TAB, CPT>(A, pb.y, (unsigned)A.nblk);
}

// The kernel of a launch: single solve or batch, preset (MSP 6 = 12 taps,
// 12 = 24 taps), tables in LDS or not, for the gridding tile of CPT cells per
// thread.  Explicitly instantiated once per CPT in persistent_tile*.hip.
template <int CPT>
const void* persistent_kernel(int MSP, bool tab, bool batch) {
  if (batch)
    return (MSP == 6) ? (tab ? (const void*)k_mfista_batch<6, true, CPT> : (const void*)k_mfista_batch<6, false, CPT>)
                      : (tab ? (const void*)k_mfista_batch<12, true, CPT> : (const void*)k_mfista_batch<12, false, CPT>);
  return (MSP == 6)
             ? (tab ? (const void*)k_mfista_persistent<6, true, CPT> : (const void*)k_mfista_persistent<6, false, CPT>)
             : (tab ? (const void*)k_mfista_persistent<12, true, CPT> : (const void*)k_mfista_persistent<12, false, CPT>);
}
extern template const void* persistent_kernel<1>(int, bool, bool);
extern template const void* persistent_kernel<2>(int, bool, bool);
extern template const void* persistent_kernel<4>(int, bool, bool);

// Gridding tiles the persistent kernels are built for: cells per thread of a
// 256-thread workgroup (CPT) and the tile's width and height in grid cells.
inline bool tile_supported(int TX, int TY) {
  return (TX == 16 && TY == 16) || (TX == 32 && TY == 16) || (TX == 32 && TY == 32);
}
}  // namespace mf
