// stages.h -- the NUFFT pipeline stages as workgroup-level device functions.
//
// Each stage processes one "task" (a group of rows, a group of visibilities,
// or a chunk of a gridding tile) with the whole workgroup, so the same code
// runs as a stand-alone kernel (task = blockIdx.x) and inside the persistent
// solver kernel (tasks dealt round-robin to the resident workgroups).  Stages
// that use LDS take the dynamic shared-memory pointer and end with a barrier,
// so a workgroup can start its next task immediately.
#pragma once

#include "nufft.h"

namespace mf {

// Forward row pass.  Task = nseq complex sequences = 2 nseq image rows from
// row0 = 2 nseq task; load(i, j) supplies the real (already E4-weighted)
// input of pixel (i, j).  The rows are placed at b = (j - shift) mod Ly,
// transformed, split into the two real rows' half spectra and written
// transposed: T[ky][i], ky in [0, Ly/2].
// a complex value stored non-temporally (the line is not kept in the
// writer's L2: less for a grid barrier's release to write back)
template <typename R>
__device__ __forceinline__ void nt_store(cplx<R>* p, const cplx<R>& v) {
  typedef R v2 __attribute__((ext_vector_type(2)));
  v2 w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<v2*>(p));
}

template <typename R, typename Load, bool NTS = false>
__device__ inline void st_rowfft_fwd(int task, int nseq, int NX, int NY, int Ly, int shift,
                                     const FFTPlan& P, const cplx<R>* __restrict__ tw,
                                     cplx<R>* __restrict__ T, cplx<R>* sm, Load load) {
  using C = cplx<R>;
  C* b0 = sm;
  C* b1 = sm + nseq * Ly;
  const int row0 = task * 2 * nseq;
  for (int t = threadIdx.x; t < nseq * Ly; t += blockDim.x) {
    const int q = t / Ly, b = t - q * Ly;
    int j = b + shift;
    if (j >= Ly) j -= Ly;
    const int i0 = row0 + 2 * q, i1 = i0 + 1;
    R re = 0, im = 0;
    if (j < NY) {
      if (i0 < NX) re = load(i0, j);
      if (i1 < NX) im = load(i1, j);
    }
    b0[t] = mk<R>(re, im);
  }
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, nseq, P, tw, -1);
  const int Lh = Ly / 2 + 1, nr = 2 * nseq;
  for (int t = threadIdx.x; t < Lh * nr; t += blockDim.x) {
    const int k = t / nr, r = t - k * nr;
    const int i = row0 + r;
    if (i >= NX) continue;
    const int q = r >> 1;
    const C z = res[q * Ly + k], zm = res[q * Ly + (k == 0 ? 0 : Ly - k)];
    C o;
    if (!(r & 1))
      o = mk<R>((R)0.5 * (z.x + zm.x), (R)0.5 * (z.y - zm.y));
    else
      o = mk<R>((R)0.5 * (z.y + zm.y), (R)-0.5 * (z.x - zm.x));
    if constexpr (NTS)
      nt_store<R>(T + (size_t)k * NX + i, o);
    else
      T[(size_t)k * NX + i] = o;
  }
  __syncthreads();
}

// Forward row pass of two images at once (the two steps of a paired
// backtracking round): sequence 0 packs rows (2 task, 2 task + 1) of image 0,
// sequence 1 the same rows of image 1, so all 256 lanes have butterflies.
// load(i, j, v0, v1) supplies pixel (i, j) of both images; the per-image
// arithmetic is that of st_rowfft_fwd.
template <typename R, typename Load2>
__device__ inline void st_rowfft_fwd_pair(int task, int NX, int NY, int Ly, int shift, const FFTPlan& P,
                                          const cplx<R>* __restrict__ tw, cplx<R>* __restrict__ T0,
                                          cplx<R>* __restrict__ T1, cplx<R>* sm, Load2 load) {
  using C = cplx<R>;
  C* b0 = sm;
  C* b1 = sm + 2 * Ly;
  const int i0 = task * 2, i1 = i0 + 1;
  for (int b = threadIdx.x; b < Ly; b += blockDim.x) {
    int j = b + shift;
    if (j >= Ly) j -= Ly;
    R re0 = 0, im0 = 0, re1 = 0, im1 = 0;
    if (j < NY) {
      if (i0 < NX) load(i0, j, re0, re1);
      if (i1 < NX) load(i1, j, im0, im1);
    }
    b0[b] = mk<R>(re0, im0);
    b0[Ly + b] = mk<R>(re1, im1);
  }
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, 2, P, tw, -1);
  const int Lh = Ly / 2 + 1;
  for (int t = threadIdx.x; t < Lh * 4; t += blockDim.x) {
    const int k = t >> 2, r = t & 3;
    const int q = r >> 1, i = i0 + (r & 1);
    if (i >= NX) continue;
    const C z = res[q * Ly + k], zm = res[q * Ly + (k == 0 ? 0 : Ly - k)];
    C o;
    if (!(r & 1))
      o = mk<R>((R)0.5 * (z.x + zm.x), (R)0.5 * (z.y - zm.y));
    else
      o = mk<R>((R)0.5 * (z.y + zm.y), (R)-0.5 * (z.x - zm.x));
    (q ? T1 : T0)[(size_t)k * NX + i] = o;
  }
  __syncthreads();
}

// Forward column pass of two transforms at once: ky row `task` of T0 -> out0
// and of T1 -> out1, NUFFT layout (mode 0 of st_colfft_fwd).
template <typename R>
__device__ inline void st_colfft_fwd_pair(int task, const cplx<R>* __restrict__ T0,
                                          const cplx<R>* __restrict__ T1, cplx<R>* __restrict__ out0,
                                          cplx<R>* __restrict__ out1, int NX, int nky, int Lx, int shift,
                                          int MSP, const FFTPlan& P, const cplx<R>* __restrict__ tw,
                                          cplx<R>* sm) {
  using C = cplx<R>;
  C* b0 = sm;
  C* b1 = sm + 2 * Lx;
  const int ky = task;
  for (int t = threadIdx.x; t < 2 * Lx; t += blockDim.x) {
    const int q = t >= Lx, a = t - q * Lx;
    int i = a + shift;
    if (i >= Lx) i -= Lx;
    b0[t] = (ky < nky && i < NX) ? (q ? T1 : T0)[(size_t)ky * NX + i] : mk<R>(0, 0);
  }
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, 2, P, tw, -1);
  const int NXh = Lx / 2;
  if (ky < nky) {
    for (int t = threadIdx.x; t < 2 * Lx; t += blockDim.x) {
      const int q = t >= Lx, k = t - q * Lx;
      C* out = q ? out1 : out0;
      const C v = res[t];
      int col = k + NXh;
      if (col >= Lx) col -= Lx;
      out[(size_t)(ky + MSP - 1) * Lx + col] = v;
      if (ky >= 1 && ky <= MSP - 1) {
        int colm = NXh - k;
        if (colm < 0) colm += Lx;
        out[(size_t)(MSP - 1 - ky) * Lx + colm] = cconj(v);
      }
    }
  }
  __syncthreads();
}

// Forward column pass: c2c along x for ky rows ky0 .. ky0+nseq-1 of T.
// mode 0: NUFFT layout Gf[ky + MSP - 1][(kx + NX) mod Lx] plus the MSP-1
//         conjugate-mirrored rows of negative ky
// mode 1: plain layout out[ky][kx]
template <typename R>
__device__ inline void st_colfft_fwd(int task, int nseq, const cplx<R>* __restrict__ T,
                                     cplx<R>* __restrict__ out, int NX, int nky, int Lx, int shift,
                                     int mode, int MSP, const FFTPlan& P,
                                     const cplx<R>* __restrict__ tw, cplx<R>* sm) {
  using C = cplx<R>;
  C* b0 = sm;
  C* b1 = sm + nseq * Lx;
  const int ky0 = task * nseq;
  for (int t = threadIdx.x; t < nseq * Lx; t += blockDim.x) {
    const int q = t / Lx, a = t - q * Lx;
    int i = a + shift;
    if (i >= Lx) i -= Lx;
    const int ky = ky0 + q;
    b0[t] = (ky < nky && i < NX) ? T[(size_t)ky * NX + i] : mk<R>(0, 0);
  }
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, nseq, P, tw, -1);
  const int NXh = Lx / 2;
  for (int t = threadIdx.x; t < nseq * Lx; t += blockDim.x) {
    const int q = t / Lx, k = t - q * Lx;
    const int ky = ky0 + q;
    if (ky >= nky) continue;
    const C v = res[t];
    if (mode == 0) {
      int col = k + NXh;
      if (col >= Lx) col -= Lx;
      out[(size_t)(ky + MSP - 1) * Lx + col] = v;
      if (ky >= 1 && ky <= MSP - 1) {
        int colm = NXh - k;
        if (colm < 0) colm += Lx;
        out[(size_t)(MSP - 1 - ky) * Lx + colm] = cconj(v);
      }
    } else {
      out[(size_t)ky * Lx + k] = v;
    }
  }
  __syncthreads();
}

// Adjoint column pass: c2c backward along x per ky row, transposed output
// U[i][ky] for the nx image rows a = (i - shift) mod Lx.
// mode 0: input in NUFFT layout Hg[ky][(kx + NX) mod Lx], cleared behind the
//         read (the gridder accumulates into zeros); mode 1: plain in[ky][kx]
template <typename R>
__device__ inline void st_colfft_inv(int task, int nseq, cplx<R>* __restrict__ in,
                                     cplx<R>* __restrict__ U, int NX, int nky, int Lx, int shift,
                                     int mode, const FFTPlan& P, const cplx<R>* __restrict__ tw,
                                     cplx<R>* sm) {
  using C = cplx<R>;
  C* b0 = sm;
  C* b1 = sm + nseq * Lx;
  const int ky0 = task * nseq;
  const int NXh = Lx / 2;
  for (int t = threadIdx.x; t < nseq * Lx; t += blockDim.x) {
    const int q = t / Lx, k = t - q * Lx;
    const int ky = ky0 + q;
    int col = k;
    if (mode == 0) {
      col = k + NXh;
      if (col >= Lx) col -= Lx;
    }
    C v = mk<R>(0, 0);
    if (ky < nky) {
      v = in[(size_t)ky * Lx + col];
      if (mode == 0) in[(size_t)ky * Lx + col] = mk<R>(0, 0);
    }
    b0[t] = v;
  }
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, nseq, P, tw, +1);
  for (int t = threadIdx.x; t < NX * nseq; t += blockDim.x) {
    const int i = t / nseq, q = t - i * nseq;
    const int ky = ky0 + q;
    if (ky >= nky) continue;
    int a = i - shift;
    if (a < 0) a += Lx;
    U[(size_t)i * nky + ky] = res[q * Lx + a];
  }
  __syncthreads();
}

// Adjoint row pass: 1-D c2r along y for two image rows per complex FFT (FFTW
// c2r semantics: only the real parts of the DC and Nyquist entries are used).
// store(i, j, v) receives the raw c2r value of pixel (i, j), j = (b+shift) mod Ly.
template <typename R, typename Store>
__device__ inline void st_rowfft_c2r(int task, int nseq, const cplx<R>* __restrict__ U, int NX,
                                     int NY, int Ly, int shift, const FFTPlan& P,
                                     const cplx<R>* __restrict__ tw, cplx<R>* sm, Store store) {
  using C = cplx<R>;
  C* b0 = sm;
  C* b1 = sm + nseq * Ly;
  const int row0 = task * 2 * nseq;
  const int Lh = Ly / 2, nky = Ly / 2 + 1;
  for (int t = threadIdx.x; t < nseq * Ly; t += blockDim.x) {
    const int q = t / Ly, b = t - q * Ly;
    const int i0 = row0 + 2 * q, i1 = i0 + 1;
    const int k = (b <= Lh) ? b : Ly - b;
    C A = (i0 < NX) ? U[(size_t)i0 * nky + k] : mk<R>(0, 0);
    C B = (i1 < NX) ? U[(size_t)i1 * nky + k] : mk<R>(0, 0);
    if (k == 0 || k == Lh) {
      A.y = 0;
      B.y = 0;
    }
    b0[t] = (b <= Lh) ? mk<R>(A.x - B.y, A.y + B.x) : mk<R>(A.x + B.y, -A.y + B.x);
  }
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, nseq, P, tw, +1);
  const int nr = 2 * nseq;
  for (int t = threadIdx.x; t < nr * NY; t += blockDim.x) {
    const int r = t / NY, j = t - r * NY;
    const int i = row0 + r;
    if (i >= NX) continue;
    int b = j - shift;
    if (b < 0) b += Ly;
    const C z = res[(r >> 1) * Ly + b];
    store(i, j, (r & 1) ? z.y : z.x);
  }
  __syncthreads();
}

// st_rowfft_c2r (one sequence) with two hooks around its input loads:
// issue() runs before them and commit() after them, before the barrier that
// publishes the LDS input, so a caller's own independent loads (issued into
// registers, stored to LDS in commit()) are in flight together with these.
template <typename R, typename Store, typename Issue, typename Commit>
__device__ inline void st_rowfft_c2r_pf(int task, const cplx<R>* __restrict__ U, int NX, int NY, int Ly, int shift,
                                        const FFTPlan& P, const cplx<R>* __restrict__ tw, cplx<R>* sm, Store store,
                                        Issue issue, Commit commit) {
  using C = cplx<R>;
  C* b0 = sm;
  C* b1 = sm + Ly;
  const int row0 = task * 2;
  const int Lh = Ly / 2, nky = Ly / 2 + 1;
  issue();
  for (int t = threadIdx.x; t < Ly; t += blockDim.x) {
    const int b = t;
    const int i0 = row0, i1 = i0 + 1;
    const int k = (b <= Lh) ? b : Ly - b;
    C A = (i0 < NX) ? U[(size_t)i0 * nky + k] : mk<R>(0, 0);
    C B = (i1 < NX) ? U[(size_t)i1 * nky + k] : mk<R>(0, 0);
    if (k == 0 || k == Lh) {
      A.y = 0;
      B.y = 0;
    }
    b0[t] = (b <= Lh) ? mk<R>(A.x - B.y, A.y + B.x) : mk<R>(A.x + B.y, -A.y + B.x);
  }
  commit();
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, 1, P, tw, +1);
  for (int t = threadIdx.x; t < 2 * NY; t += blockDim.x) {
    const int r = t / NY, j = t - r * NY;
    const int i = row0 + r;
    if (i >= NX) continue;
    int b = j - shift;
    if (b < 0) b += Ly;
    const C z = res[b];
    store(i, j, (r & 1) ? z.y : z.x);
  }
  __syncthreads();
}

// Tap gather of the visibilities of one task, G = 16 (MSP 6) or 32 (MSP 12)
// lanes per visibility and blockDim / G visibilities per task; lane l owns tap
// column kx = mx + l.  The tap count 2 MSP is a compile-time constant so the
// 2 MSP gathers of a lane are issued together.  Returns true on the group
// leader of a valid visibility s, with F = (A x)_s in the reflected frame
// (the NU_SIGN conjugation and 1/(4 NX NY) applied).
template <int MSPC>
struct DegridShape {
  static constexpr int S = 2 * MSPC;
  static constexpr int G = (S <= 16) ? 16 : 32;
};

template <typename R, int MSPC>
__device__ inline bool degrid_core(int task, const GridGeom& g, const VisParams<R>& vp,
                                   const cplx<R>* __restrict__ Gf, int& s, cplx<R>& F) {
  using C = cplx<R>;
  constexpr int S = DegridShape<MSPC>::S, G = DegridShape<MSPC>::G;
  const int gid = task * blockDim.x + threadIdx.x;
  s = gid / G;
  const int lane = threadIdx.x & (G - 1);
  const R t3x = (R)g.t3x, t3y = (R)g.t3y;
  C acc = mk<R>(0, 0);
  // lane b evaluates the row weight of tap row b once; the lanes of the group
  // read it by shuffle (the weights depend on the row, not on the lane's
  // column).  Inside the branch every source lane b < S of an active lane's
  // group is active too (the group shares one visibility).
  const C a = (s < g.M) ? GP(vp.a)[s] : mk<R>(0, 0);
  const R l = (R)(lane - MSPC + 1);
  const R gyl = dev_exp(l * a.y - l * l * t3y);
  if (s < g.M && lane < S) {
    const int2 m = GP(vp.mxy)[s];
    const R gx = dev_exp(l * a.x - l * l * t3x);
    const C* col = Gf + (size_t)m.y * g.Lx + (m.x + g.NX + lane - MSPC + 1);
    C gv[S];
#pragma unroll
    for (int b = 0; b < S; b++) gv[b] = col[(size_t)b * g.Lx];
#pragma unroll
    for (int b = 0; b < S; b++) {
      const R gy = __shfl(gyl, b, G);
      acc.x += gy * gv[b].x;
      acc.y += gy * gv[b].y;
    }
    acc.x *= gx;
    acc.y *= gx;
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
    acc.x += __shfl_xor(acc.x, off, 64);
    acc.y += __shfl_xor(acc.y, off, 64);
  }
  if (s < g.M && lane == 0) {
    const R e1s = GP(vp.e1)[s] * (R)g.inv_MM;
    F = mk<R>(e1s * acc.x, -e1s * acc.y);
    return true;
  }
  return false;
}

// Degridding of one task.
//   mode 0: r = (V' - A x) w in R (stored to rz and/or rz32); v[0] = |r|^2
//   mode 1: x is a step d: t = (A d) w, r_new = rz64 - t; v[0] = |r_new|^2,
//           v[1] = |t|^2
// v[] holds this thread's contribution (nonzero on the group leaders only).
template <typename R, int MSPC>
__device__ inline void st_degrid(int task, const GridGeom& g, const VisParams<R>& vp,
                                 const cplx<R>* __restrict__ Gf, cplx<R>* __restrict__ rz,
                                 float2* __restrict__ rz32, const double2* __restrict__ rz64,
                                 int mode, cplx<R>* __restrict__ model,
                                 const int* __restrict__ perm, const int* __restrict__ yneg,
                                 double v[2]) {
  using C = cplx<R>;
  int s;
  C F;
  const bool lead = degrid_core<R, MSPC>(task, g, vp, Gf, s, F);
  v[0] = 0.0;
  v[1] = 0.0;
  if (lead) {
    const R w = GP(vp.w)[s];
    if (mode == 0) {
      const C V = GP(vp.vis)[s];
      const C r = mk<R>((V.x - F.x) * w, (V.y - F.y) * w);
      v[0] = (double)r.x * (double)r.x + (double)r.y * (double)r.y;
      if (rz) rz[s] = r;
      if (rz32) rz32[s] = make_float2((float)r.x, (float)r.y);
    } else {
      const double tx = (double)F.x * (double)w, ty = (double)F.y * (double)w;
      const double2 r0 = rz64[s];
      const double rx = r0.x - tx, ry = r0.y - ty;
      v[0] = rx * rx + ry * ry;
      v[1] = tx * tx + ty * ty;
    }
    if (model) model[perm[s]] = yneg[s] ? cconj(F) : F;
  }
}

// FP64 degridding into weighted model vectors m = W A x (reflected frame).
//   init (mz_in == nullptr): t = W A x; out0 = out1 = t; v[0] += |W V' - t|^2
//   step (mz_in != nullptr): t = W A d; out0 = mz + t (the model of z + d);
//        v[0] += |W V' - mz - t|^2, v[1] += |t|^2
template <int MSPC>
__device__ inline void st_degrid_model(int task, const GridGeom& g, const VisParams<double>& vp,
                                       const double2* __restrict__ Gf,
                                       const double2* __restrict__ mz_in,
                                       double2* __restrict__ out0, double2* __restrict__ out1,
                                       double v[2]) {
  int s;
  double2 F;
  const bool lead = degrid_core<double, MSPC>(task, g, vp, Gf, s, F);
  if (lead) {
    const double w = GP(vp.w)[s];
    const double2 V = GP(vp.vis)[s];
    const double tx = F.x * w, ty = F.y * w;
    if (!mz_in) {
      const double rx = V.x * w - tx, ry = V.y * w - ty;
      v[0] += rx * rx + ry * ry;
      out0[s] = make_double2(tx, ty);
      out1[s] = make_double2(tx, ty);
    } else {
      const double2 mz = mz_in[s];
      const double rx = V.x * w - mz.x - tx, ry = V.y * w - mz.y - ty;
      v[0] += rx * rx + ry * ry;
      v[1] += tx * tx + ty * ty;
      out0[s] = make_double2(mz.x + tx, mz.y + ty);
    }
  }
}

// The constants of one visibility that the degridder needs: tap slopes,
// grid cell, E1 / (4 NX NY), weight and data (reflected frame).  They do not
// change during a solve, so the persistent kernel caches the records of its
// workgroup's degrid tasks in LDS once per launch (dvis_load for the rest).
struct alignas(16) DVis {
  double ax, ay, e1s, w, vx, vy;
  int mx, my;
};

__device__ inline DVis dvis_load(const VisParams<double>& vp, const GridGeom& g, int s) {
  DVis r{};
  if (s < g.M) {
    const double2 a = gld(GP(vp.a) + s), V = gld(GP(vp.vis) + s);
    const int2 m = gld(GP(vp.mxy) + s);
    r.ax = a.x;
    r.ay = a.y;
    r.e1s = gld(GP(vp.e1) + s) * g.inv_MM;
    r.w = gld(GP(vp.w) + s);
    r.vx = V.x;
    r.vy = V.y;
    r.mx = m.x;
    r.my = m.y;
  }
  return r;
}

// Lane layout of the paired degrid: a group of GL lanes per visibility (lane
// li < 2 MSP owns tap column li), VPW groups per 64-wide wave, VPT
// visibilities per 256-thread task.  12 taps use groups of 12 lanes, five to a
// wave (60 of 64 lanes busy): a task then covers 20 visibilities instead of 16,
// and no workgroup runs more than two of these latency-bound tasks at 20000
// visibilities on 512 workgroups.
template <int MSPC>
struct PairShape {
  static constexpr int GL = (2 * MSPC <= 12) ? 12 : 32;
  static constexpr int VPW = 64 / GL;
  static constexpr int VPT = VPW * (MF_BLOCK / 64);
};

// Tap gather of two spectra with the same taps (degrid_core on Gf0 and Gf1)
// for visibility s with constants r; lane li of group gi of its wave.  The
// tap rows are loaded in groups of up to 12 (both spectra, 4 VGPRs per row
// and spectrum).  For the 24-tap kernel a scheduling barrier keeps each
// group's loads ahead of its FMAs: the gather is latency-bound, and left to
// itself the scheduler interleaved loads and FMAs, which cost that stage
// ~30 %.  The 12-tap kernel stays unpinned: there the barrier's extra live
// registers cost the kernel its second workgroup per CU (measured on the
// MI355X).
template <int MSPC>
__device__ inline bool degrid_core_pair(int s, bool valid, int gi, int li, const DVis& r, const GridGeom& g,
                                        const double2* __restrict__ Gf0, const double2* __restrict__ Gf1,
                                        double2& F0, double2& F1) {
  constexpr int S = 2 * MSPC, GL = PairShape<MSPC>::GL;
  constexpr int RG = (S <= 12) ? S : 12;  // rows per load group
  static_assert(S % RG == 0, "tap rows split into whole groups");
  const int base = gi * GL;  // wave lane of the group's lane 0
  const double t3x = g.t3x, t3y = g.t3y;
  double2 a0 = make_double2(0, 0), a1 = make_double2(0, 0);
  const double l = (double)(li - MSPC + 1);
  const double gyl = dev_exp(l * r.ay - l * l * t3y);  // row weight of tap row `li`, shared by shuffle
  if (valid && li < S) {
    const double gx = dev_exp(l * r.ax - l * l * t3x);
    const size_t off = (size_t)r.my * g.Lx + (r.mx + g.NX + li - MSPC + 1);
    const double2* c0 = Gf0 + off;
    const double2* c1 = Gf1 + off;
    if constexpr (MSPC > 8) {
#pragma unroll
      for (int b0 = 0; b0 < S; b0 += RG) {
        double2 u0[RG], u1[RG];
#pragma unroll
        for (int b = 0; b < RG; b++) {
          u0[b] = gld(c0 + (size_t)(b0 + b) * g.Lx);
          u1[b] = gld(c1 + (size_t)(b0 + b) * g.Lx);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int b = 0; b < RG; b++) {
          const double gy = __shfl(gyl, base + b0 + b, 64);
          a0.x += gy * u0[b].x;
          a0.y += gy * u0[b].y;
          a1.x += gy * u1[b].x;
          a1.y += gy * u1[b].y;
        }
      }
    } else {
#pragma unroll
      for (int b = 0; b < S; b++) {
        const double2 u0 = gld(c0 + (size_t)b * g.Lx), u1 = gld(c1 + (size_t)b * g.Lx);
        const double gy = __shfl(gyl, base + b, 64);
        a0.x += gy * u0.x;
        a0.y += gy * u0.y;
        a1.x += gy * u1.x;
        a1.y += gy * u1.y;
      }
    }
    a0.x *= gx;
    a0.y *= gx;
    a1.x *= gx;
    a1.y *= gx;
  }
  if constexpr ((GL & (GL - 1)) == 0) {
#pragma unroll
    for (int off = GL / 2; off > 0; off >>= 1) {
      a0.x += __shfl_xor(a0.x, off, 64);
      a0.y += __shfl_xor(a0.y, off, 64);
      a1.x += __shfl_xor(a1.x, off, 64);
      a1.y += __shfl_xor(a1.y, off, 64);
    }
  } else {
    // shuffle-down tree over the GL lanes of the group: lane 0 ends with the
    // sum (a lane adds its partner li + off only while that is in the group)
    const int wl = base + li;
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) {
      const bool in = li + off < GL;
      const double p0 = __shfl(a0.x, wl + off, 64), p1 = __shfl(a0.y, wl + off, 64);
      const double p2 = __shfl(a1.x, wl + off, 64), p3 = __shfl(a1.y, wl + off, 64);
      if (in) {
        a0.x += p0;
        a0.y += p1;
        a1.x += p2;
        a1.y += p3;
      }
    }
    static_assert(GL <= 16, "tree covers up to 16 lanes");
  }
  if (valid && li == 0) {
    F0 = make_double2(r.e1s * a0.x, -r.e1s * a0.y);
    F1 = make_double2(r.e1s * a1.x, -r.e1s * a1.y);
    return true;
  }
  return false;
}

// Step mode of st_degrid_model for the two steps of a paired round:
// out_k = mz + W A d_k; v[2k] += |W V' - mz - W A d_k|^2, v[2k+1] += |W A d_k|^2.
// Task `task` covers visibilities task VPT .. task VPT + VPT - 1 (PairShape).
// dv: the task's visibility records in LDS, or nullptr to load them.  The
// group leader's model load is issued before the tap gather.
template <int MSPC>
__device__ inline void st_degrid_model_pair(int task, const GridGeom& g, const VisParams<double>& vp,
                                            const DVis* dv, const double2* __restrict__ Gf0,
                                            const double2* __restrict__ Gf1, const double2* __restrict__ mz_in,
                                            double2* __restrict__ out0, double2* __restrict__ out1, double v[4]) {
  using PS = PairShape<MSPC>;
  const int wl = threadIdx.x & 63;
  const int gi = wl / PS::GL, li = wl - gi * PS::GL;
  const int j = (threadIdx.x >> 6) * PS::VPW + gi;  // visibility of this lane within the task
  const int s = task * PS::VPT + j;
  const bool valid = gi < PS::VPW && s < g.M;
  const bool lead = valid && li == 0;
  double2 mz = make_double2(0, 0);
  if (lead) mz = gld(mz_in + s);
  DVis r{};
  if (gi < PS::VPW) r = dv ? dv[j] : dvis_load(vp, g, s);
  double2 F0, F1;
  if (degrid_core_pair<MSPC>(s, valid, gi, li, r, g, Gf0, Gf1, F0, F1)) {
    const double w = r.w;
    const double bx = r.vx * w - mz.x, by = r.vy * w - mz.y;
    {
      const double tx = F0.x * w, ty = F0.y * w;
      const double rx = bx - tx, ry = by - ty;
      v[0] += rx * rx + ry * ry;
      v[1] += tx * tx + ty * ty;
      gst(out0 + s, make_double2(mz.x + tx, mz.y + ty));
    }
    {
      const double tx = F1.x * w, ty = F1.y * w;
      const double rx = bx - tx, ry = by - ty;
      v[2] += rx * rx + ry * ry;
      v[3] += tx * tx + ty * ty;
      gst(out1 + s, make_double2(mz.x + tx, mz.y + ty));
    }
  }
}

// Gridding of one chunk of a tile's entry list.  Each of the 4 wavefronts
// accumulates into its own LDS copy of the TY x TX tile (its lanes own
// distinct taps, so no LDS atomics); the copies are summed and every touched
// cell is added to Hg with one global atomic.  Deposit of visibility s:
// E1 w conj(rz) at the taps with ky >= 0, and its conjugate at -k for the
// taps with ky <= 0 (mirror entries), which builds S = G + conj G(-k).
// resid(s) returns the weighted residual r_s of visibility s.
template <typename R, typename Resid>
__device__ inline void st_grid(int chunk, const GridGeom& g, const VisParams<R>& vp, Resid resid,
                               const int4* __restrict__ chunks, const unsigned* __restrict__ ent,
                               cplx<R>* __restrict__ Hg, cplx<R>* sm) {
  using C = cplx<R>;
  C* acc = sm;
  const int TX = g.TX, TY = g.TY, TT = TX * TY;
  const int4 ch = chunks[chunk];
  const int tile = ch.x;
  const int tr = tile / g.ntx, tc = tile - tr * g.ntx;
  const int rb = tr * TY, cb = tc * TX;
  for (int t = threadIdx.x; t < 4 * TT; t += blockDim.x) acc[t] = mk<R>(0, 0);
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  C* my = acc + wave * TT;
  const int MSP = g.MSP, NX = g.NX;
  const R t3x = (R)g.t3x, t3y = (R)g.t3y;
  for (int e = ch.y + wave; e < ch.z; e += 4) {
    const unsigned en = ent[e];
    const int s = (int)(en & 0x7fffffffu);
    const bool mir = (en >> 31) != 0;
    const int2 m = GP(vp.mxy)[s];
    const C a = GP(vp.a)[s];
    const R ew = GP(vp.e1)[s] * GP(vp.w)[s];
    const C r = resid(s);
    C d = mk<R>(ew * r.x, -ew * r.y);
    int lp0, lp1, l0, l1;
    if (!mir) {
      lp0 = max(-MSP + 1, rb - m.y);
      lp1 = min(MSP + 1, rb + TY - m.y);
      l0 = max(-MSP + 1, cb - m.x - NX);
      l1 = min(MSP + 1, cb + TX - m.x - NX);
    } else {
      d.y = -d.y;
      lp0 = max(-MSP + 1, -m.y - rb - TY + 1);
      lp1 = min(MSP + 1, -m.y - rb + 1);
      l0 = max(-MSP + 1, NX - m.x - cb - TX + 1);
      l1 = min(MSP + 1, NX - m.x - cb + 1);
    }
    const int nl = l1 - l0, nlp = lp1 - lp0;
    if (nl <= 0 || nlp <= 0) continue;
    const int ntap = nl * nlp;
    for (int t = lane; t < ntap; t += 64) {
      const int ia = t / nl, ib = t - ia * nl;
      const int lp = lp0 + ia, l = l0 + ib;
      const R fl = (R)l, flp = (R)lp;
      const R wgt = dev_exp(fl * a.x - fl * fl * t3x + flp * a.y - flp * flp * t3y);
      int row, col;
      if (!mir) {
        row = m.y + lp - rb;
        col = m.x + NX + l - cb;
      } else {
        row = -(m.y + lp) - rb;
        col = NX - m.x - l - cb;
      }
      C& c = my[row * TX + col];
      c.x += wgt * d.x;
      c.y += wgt * d.y;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < TT; t += blockDim.x) {
    const int rr = t / TX, cc = t - rr * TX;
    const int row = rb + rr, col = cb + cc;
    C sum = acc[t];
    sum = cadd(sum, acc[TT + t]);
    sum = cadd(sum, acc[2 * TT + t]);
    sum = cadd(sum, acc[3 * TT + t]);
    if (row <= g.NY && col < g.Lx && (sum.x != (R)0 || sum.y != (R)0)) {
      R* dst = reinterpret_cast<R*>(Hg + (size_t)row * g.Lx + col);
      unsafeAtomicAdd(dst, sum.x);
      unsafeAtomicAdd(dst + 1, sum.y);
    }
  }
  __syncthreads();
}

}  // namespace mf
