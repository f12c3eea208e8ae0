// nufft.hip -- Gaussian-gridding NUFFT kernels for gfx950.
//
// Forward A.x  (reference NUFFT2d2, c/mfista_nufft_lib.c:187-243 and
//               c++/mfista_nufft_lib.cpp:440-504):
//   k_rowfft_fwd  E4-weighted, centred rows -> r2c along y (two rows per complex
//                 FFT) -> T[ky][i]  (transposed through LDS)
//   k_colfft_fwd  c2c along x per ky -> Gf[ky][kx] (fft-shifted columns, plus
//                 MSP-1 conjugate-mirrored rows of negative ky)
//   k_degrid      G lanes per visibility gather the (2MSP)^2 taps, then the
//                 weighted residual and its FP64 chi^2 partial are fused in
// Adjoint Re(A^H r) (reference NUFFT2d1, c/mfista_nufft_lib.c:123-185 and
//               c++/mfista_nufft_lib.cpp:189-438):
//   k_grid        one workgroup per chunk of a tile's visibility list; each of
//                 the 4 wavefronts accumulates into its own LDS copy (lanes own
//                 distinct taps, so no LDS atomics), then one global atomic
//                 add per touched cell
//   k_colfft_inv  c2c backward along x per ky -> U[i][ky]
//   k_rowfft_c2r  c2r along y (FFTW semantics), crop, * E4 / (8 NX NY)
#include "nufft.h"
#include "stages.h"

namespace mf {

// --------------------------------------------------------------------------
// stand-alone kernels: one stage task per workgroup (stages.h)
template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_rowfft_fwd(const R* __restrict__ x, const R* __restrict__ e4x, const R* __restrict__ e4y,
             cplx<R>* __restrict__ T, int NX, int NY, int Ly, int shift, FFTPlan P,
             const cplx<R>* __restrict__ tw, int nseq, const int* skip) {
  if (skip && *skip) return;
  extern __shared__ double4 smem_d4[];
  auto load = [&](int i, int j) -> R {
    return x[(size_t)i * NY + j] * (e4x ? e4x[i] : (R)1) * (e4y ? e4y[j] : (R)1);
  };
  st_rowfft_fwd<R>(blockIdx.x, nseq, NX, NY, Ly, shift, P, tw, T,
                   reinterpret_cast<cplx<R>*>(smem_d4), load);
}

template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_colfft_fwd(const cplx<R>* __restrict__ T, cplx<R>* __restrict__ out, int NX, int nky, int Lx,
             int shift, int mode, int MSP, FFTPlan P, const cplx<R>* __restrict__ tw, int nseq,
             const int* skip) {
  if (skip && *skip) return;
  extern __shared__ double4 smem_d4[];
  st_colfft_fwd<R>(blockIdx.x, nseq, T, out, NX, nky, Lx, shift, mode, MSP, P, tw,
                   reinterpret_cast<cplx<R>*>(smem_d4));
}

template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_colfft_inv(cplx<R>* __restrict__ in, cplx<R>* __restrict__ U, int NX, int nky, int Lx,
             int shift, int mode, FFTPlan P, const cplx<R>* __restrict__ tw, int nseq,
             const int* skip) {
  if (skip && *skip) return;
  extern __shared__ double4 smem_d4[];
  st_colfft_inv<R>(blockIdx.x, nseq, in, U, NX, nky, Lx, shift, mode, P, tw,
                   reinterpret_cast<cplx<R>*>(smem_d4));
}

template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_rowfft_c2r(const cplx<R>* __restrict__ U, const R* __restrict__ e4x, const R* __restrict__ e4y,
             R scale, R* __restrict__ g, int NX, int NY, int Ly, int shift, FFTPlan P,
             const cplx<R>* __restrict__ tw, int nseq, const int* skip) {
  if (skip && *skip) return;
  extern __shared__ double4 smem_d4[];
  auto store = [&](int i, int j, R v) {
    v *= scale;
    if (e4x) v *= e4x[i];
    if (e4y) v *= e4y[j];
    g[(size_t)i * NY + j] = v;
  };
  st_rowfft_c2r<R>(blockIdx.x, nseq, U, NX, NY, Ly, shift, P, tw,
                   reinterpret_cast<cplx<R>*>(smem_d4), store);
}

template <typename R, int MSPC>
__global__ void __launch_bounds__(MF_BLOCK)
k_degrid(GridGeom g, VisParams<R> vp, const cplx<R>* __restrict__ Gf, cplx<R>* __restrict__ rz,
         float2* __restrict__ rz32, const double2* __restrict__ rz64, double* __restrict__ chi_part,
         double* __restrict__ lip_part, int mode, cplx<R>* __restrict__ model,
         const int* __restrict__ perm, const int* __restrict__ yneg, const int* skip) {
  if (skip && *skip) return;
  __shared__ double red[32];
  double v[2];
  st_degrid<R, MSPC>(blockIdx.x, g, vp, Gf, rz, rz32, rz64, mode, model, perm, yneg, v);
  if (chi_part) {
    block_sum<2>(v, red);
    if (threadIdx.x == 0) {
      chi_part[blockIdx.x] = v[0];
      if (lip_part) lip_part[blockIdx.x] = v[1];
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_grid(GridGeom g, VisParams<R> vp, const cplx<R>* __restrict__ rz, const int4* __restrict__ chunks,
       const unsigned* __restrict__ ent, cplx<R>* __restrict__ Hg, const int* skip) {
  if (skip && *skip) return;
  extern __shared__ double4 smem_d4[];
  st_grid<R>(blockIdx.x, g, vp, [&](int s) { return rz[s]; }, chunks, ent, Hg,
             reinterpret_cast<cplx<R>*>(smem_d4));
}

// --------------------------------------------------------------------------
// setup: grid cell of every visibility (FP64 maths, as preNUFFT computes it)
__device__ inline void vis_cell(double u, double v, int NX, int NY, int preset, bool& flip,
                                double& uu, double& vv, int& mx, int& my) {
  const double pi = 3.14159265358979323846;
  flip = v < 0;
  uu = flip ? -u : u;
  vv = flip ? -v : v;
  const double fx = uu * (2.0 * NX) / (2.0 * pi), fy = vv * (2.0 * NY) / (2.0 * pi);
  if (preset == 0) {
    mx = (int)round(fx);
    my = (int)round(fy);
  } else {
    mx = (int)floor(fx);
    my = (int)floor(fy);
  }
}

__global__ void k_vis_keys(const double* __restrict__ u, const double* __restrict__ v, int M,
                           int NX, int NY, int preset, int MSP, int TX, int TY, int ntx,
                           int* __restrict__ keys, int* __restrict__ idx, int* err) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  bool flip;
  double uu, vv;
  int mx, my;
  vis_cell(u[k], v[k], NX, NY, preset, flip, uu, vv, mx, my);
  const bool ok = (mx >= -NX + MSP) && (mx <= NX - MSP - 1) && (my >= -NY + MSP) &&
                  (my <= NY - MSP - 1) && isfinite(u[k]) && isfinite(v[k]);
  if (!ok) atomicOr(err, 1);
  int tr = ok ? max(my, 0) / TY : 0, tc = ok ? (mx + NX) / TX : 0;
  keys[k] = tr * ntx + tc;
  idx[k] = k;
}

__global__ void k_vis_params(const int* __restrict__ perm, const double* __restrict__ u,
                             const double* __restrict__ v, const double* __restrict__ vr,
                             const double* __restrict__ vi, const double* __restrict__ sd, int M,
                             int NX, int NY, int preset, int MSP, VisParams<float> pf,
                             VisParams<double> pd, int* __restrict__ yneg) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  const int k = perm[s];
  const double pi = 3.14159265358979323846;
  bool flip;
  double uu, vv;
  int mx, my;
  vis_cell(u[k], v[k], NX, NY, preset, flip, uu, vv, mx, my);
  const double Mrx = 2.0 * NX, Mry = 2.0 * NY;
  double taux, tauy;
  if (preset == 0) {
    taux = 12.0 / ((double)NX * NX);
    tauy = 12.0 / ((double)NY * NY);
  } else {
    taux = (double)MSP / ((double)NX * NX);
    tauy = taux;  // the C++ preset uses one tau for both axes
  }
  const double xix = 2.0 * pi * mx / Mrx, xiy = 2.0 * pi * my / Mry;
  const double dx = uu - xix, dy = vv - xiy;
  const double e1 = exp(-dx * dx / (4.0 * taux) - dy * dy / (4.0 * tauy));
  const double ax = pi * dx / (Mrx * taux), ay = pi * dy / (Mry * tauy);
  const double Vr = vr[k], Vi = flip ? -vi[k] : vi[k];
  const double w = 1.0 / sd[k];
  pd.mxy[s] = make_int2(mx, my);
  pd.a[s] = make_double2(ax, ay);
  pd.e1[s] = e1;
  pd.vis[s] = make_double2(Vr, Vi);
  pd.w[s] = w;
  pf.mxy[s] = make_int2(mx, my);
  pf.a[s] = make_float2((float)ax, (float)ay);
  pf.e1[s] = (float)e1;
  pf.vis[s] = make_float2((float)Vr, (float)Vi);
  pf.w[s] = (float)w;
  yneg[s] = flip ? 1 : 0;
}

// up to 6 gridding-tile entries per visibility: <= 4 for the direct footprint
// (clipped to ky >= 0) and <= 2 for the mirrored part of taps at ky <= 0.
__global__ void k_tile_entries(const int2* __restrict__ mxy, int M, int NX, int NY, int MSP,
                               int TX, int TY, int ntx, int nty, int slots, int* __restrict__ keys,
                               unsigned* __restrict__ vals) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  const int2 m = mxy[s];
  const int sentinel = ntx * nty;
  int n = 0;
  int* kk = keys + (size_t)slots * s;
  unsigned* vv = vals + (size_t)slots * s;
  {
    const int r0 = max(0, m.y - MSP + 1), r1 = m.y + MSP;
    const int c0 = m.x + NX - MSP + 1, c1 = m.x + NX + MSP;
    for (int tr = r0 / TY; tr <= r1 / TY && tr < nty; tr++)
      for (int tc = c0 / TX; tc <= c1 / TX && tc < ntx; tc++)
        if (n < slots) {
          kk[n] = tr * ntx + tc;
          vv[n] = (unsigned)s;
          n++;
        }
  }
  if (m.y - MSP + 1 <= 0) {
    const int r1 = MSP - 1 - m.y;
    const int c0 = NX - m.x - MSP, c1 = NX - m.x + MSP - 1;
    for (int tr = 0; tr <= r1 / TY && tr < nty; tr++)
      for (int tc = c0 / TX; tc <= c1 / TX && tc < ntx; tc++)
        if (n < slots) {
          kk[n] = tr * ntx + tc;
          vv[n] = (unsigned)s | 0x80000000u;
          n++;
        }
  }
  for (; n < slots; n++) {
    kk[n] = sentinel;
    vv[n] = 0;
  }
}

__global__ void k_tile_ranges(const int* __restrict__ keys, int n, int ntiles,
                              int* __restrict__ start, int* __restrict__ end) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int k = keys[e];
  if (k >= ntiles) return;
  if (e == 0 || keys[e - 1] != k) start[k] = e;
  if (e == n - 1 || keys[e + 1] != k) end[k] = e + 1;
}

// --------------------------------------------------------------------------
// sequences per workgroup: bounded by 64 KiB of LDS (two buffers) and chosen
// so that a pass launches about 256 workgroups or more (one per CU)
static inline int fft_nseq(int L, size_t csize, int nseq_total) {
  int n = (int)(65536 / (2 * (size_t)L * csize));
  if (n < 1) n = 1;
  int want = nseq_total / 256;
  if (want < 1) want = 1;
  if (n > want) n = want;
  if (n > 8) n = 8;
  return n;
}

template <typename R>
void launch_rowfft_fwd(const R* x, const R* e4x, const R* e4y, cplx<R>* T, int NX, int NY, int Ly,
                       int shift, const FFTPlan& P, const cplx<R>* tw, hipStream_t st,
                       const int* skip) {
  const int nseq = fft_nseq(Ly, sizeof(cplx<R>), (NX + 1) / 2);
  const int nblk = (NX + 2 * nseq - 1) / (2 * nseq);
  const size_t sh = 2 * (size_t)nseq * Ly * sizeof(cplx<R>);
  hipLaunchKernelGGL(k_rowfft_fwd<R>, dim3(nblk), dim3(MF_BLOCK), sh, st, x, e4x, e4y, T, NX, NY,
                     Ly, shift, P, tw, nseq, skip);
}

template <typename R>
void launch_colfft_fwd(const cplx<R>* T, cplx<R>* out, int NX, int nky, int Lx, int shift,
                       int mode, int MSP, const FFTPlan& P, const cplx<R>* tw, hipStream_t st,
                       const int* skip) {
  const int nseq = fft_nseq(Lx, sizeof(cplx<R>), nky);
  const int nblk = (nky + nseq - 1) / nseq;
  const size_t sh = 2 * (size_t)nseq * Lx * sizeof(cplx<R>);
  hipLaunchKernelGGL(k_colfft_fwd<R>, dim3(nblk), dim3(MF_BLOCK), sh, st, T, out, NX, nky, Lx,
                     shift, mode, MSP, P, tw, nseq, skip);
}

template <typename R>
void launch_colfft_inv(cplx<R>* in, cplx<R>* U, int NX, int nky, int Lx, int shift,
                       int mode, const FFTPlan& P, const cplx<R>* tw, hipStream_t st,
                       const int* skip) {
  const int nseq = fft_nseq(Lx, sizeof(cplx<R>), nky);
  const int nblk = (nky + nseq - 1) / nseq;
  const size_t sh = 2 * (size_t)nseq * Lx * sizeof(cplx<R>);
  hipLaunchKernelGGL(k_colfft_inv<R>, dim3(nblk), dim3(MF_BLOCK), sh, st, in, U, NX, nky, Lx,
                     shift, mode, P, tw, nseq, skip);
}

template <typename R>
void launch_rowfft_c2r(const cplx<R>* U, const R* e4x, const R* e4y, R scale, R* g, int NX,
                       int NY, int Ly, int shift, const FFTPlan& P, const cplx<R>* tw,
                       hipStream_t st, const int* skip) {
  const int nseq = fft_nseq(Ly, sizeof(cplx<R>), (NX + 1) / 2);
  const int nblk = (NX + 2 * nseq - 1) / (2 * nseq);
  const size_t sh = 2 * (size_t)nseq * Ly * sizeof(cplx<R>);
  hipLaunchKernelGGL(k_rowfft_c2r<R>, dim3(nblk), dim3(MF_BLOCK), sh, st, U, e4x, e4y, scale, g,
                     NX, NY, Ly, shift, P, tw, nseq, skip);
}

// the presets use MSP 6 (CXX-12) or 12 (C-24); the degridder is compiled for both
static inline int degrid_lanes(int MSP) { return (MSP == 6) ? DegridShape<6>::G : DegridShape<12>::G; }

int degrid_blocks(int M, int MSP) {
  const int G = degrid_lanes(MSP);
  const int per = MF_BLOCK / G;
  return (M + per - 1) / per;
}

template <typename R>
void launch_degrid(const GridGeom& g, const VisParams<R>& vp, const cplx<R>* Gf, cplx<R>* rz,
                   float2* rz32, const double2* rz64, double* chi_part, double* lip_part, int mode,
                   int nblk, cplx<R>* model, const int* perm, const int* yneg_sorted,
                   hipStream_t st, const int* skip) {
  if (g.MSP == 6)
    hipLaunchKernelGGL((k_degrid<R, 6>), dim3(nblk), dim3(MF_BLOCK), 0, st, g, vp, Gf, rz, rz32, rz64,
                       chi_part, lip_part, mode, model, perm, yneg_sorted, skip);
  else
    hipLaunchKernelGGL((k_degrid<R, 12>), dim3(nblk), dim3(MF_BLOCK), 0, st, g, vp, Gf, rz, rz32, rz64,
                       chi_part, lip_part, mode, model, perm, yneg_sorted, skip);
}

template <typename R>
void launch_grid(const GridGeom& g, const VisParams<R>& vp, const cplx<R>* rz, const int4* chunks,
                 int nchunks, const unsigned* entries, cplx<R>* Hg, hipStream_t st,
                 const int* skip) {
  if (nchunks <= 0) return;
  const size_t sh = 4 * (size_t)g.TX * g.TY * sizeof(cplx<R>);
  hipLaunchKernelGGL(k_grid<R>, dim3(nchunks), dim3(MF_BLOCK), sh, st, g, vp, rz, chunks, entries,
                     Hg, skip);
}

void launch_vis_keys(const double* u, const double* v, int M, int NX, int NY, int preset,
                     int MSP, int TX, int TY, int ntx, int* keys, int* idx, int* err,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_vis_keys, dim3((M + 255) / 256), dim3(256), 0, st, u, v, M, NX, NY, preset,
                     MSP, TX, TY, ntx, keys, idx, err);
}

void launch_vis_params(const int* perm, const double* u, const double* v, const double* vr,
                       const double* vi, const double* sd, int M, int NX, int NY, int preset,
                       int MSP, VisParams<float> pf, VisParams<double> pd, int* yneg,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_vis_params, dim3((M + 255) / 256), dim3(256), 0, st, perm, u, v, vr, vi, sd,
                     M, NX, NY, preset, MSP, pf, pd, yneg);
}

void launch_tile_entries(const int2* mxy, int M, int NX, int NY, int MSP, int TX, int TY,
                         int ntx, int nty, int slots, int* keys, unsigned* vals, hipStream_t st) {
  hipLaunchKernelGGL(k_tile_entries, dim3((M + 255) / 256), dim3(256), 0, st, mxy, M, NX, NY, MSP,
                     TX, TY, ntx, nty, slots, keys, vals);
}

void launch_tile_ranges(const int* keys_sorted, int n, int ntiles, int* start, int* end,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_tile_ranges, dim3((n + 255) / 256), dim3(256), 0, st, keys_sorted, n,
                     ntiles, start, end);
}

#define MF_INST(R)                                                                               \
  template void launch_rowfft_fwd<R>(const R*, const R*, const R*, cplx<R>*, int, int, int, int, \
                                     const FFTPlan&, const cplx<R>*, hipStream_t, const int*);   \
  template void launch_colfft_fwd<R>(const cplx<R>*, cplx<R>*, int, int, int, int, int, int,     \
                                     const FFTPlan&, const cplx<R>*, hipStream_t, const int*);   \
  template void launch_colfft_inv<R>(cplx<R>*, cplx<R>*, int, int, int, int, int,                \
                                     const FFTPlan&, const cplx<R>*, hipStream_t, const int*);   \
  template void launch_rowfft_c2r<R>(const cplx<R>*, const R*, const R*, R, R*, int, int, int,   \
                                     int, const FFTPlan&, const cplx<R>*, hipStream_t,           \
                                     const int*);                                                \
  template void launch_degrid<R>(const GridGeom&, const VisParams<R>&, const cplx<R>*, cplx<R>*, \
                                 float2*, const double2*, double*, double*, int, int, cplx<R>*,  \
                                 const int*, const int*, hipStream_t, const int*);               \
  template void launch_grid<R>(const GridGeom&, const VisParams<R>&, const cplx<R>*, const int4*, \
                               int, const unsigned*, cplx<R>*, hipStream_t, const int*);

MF_INST(float)
MF_INST(double)

}  // namespace mf
