// nufft.h -- device data layout and launchers of the Gaussian-gridding NUFFT.
//
// Maths (Greengard & Lee 2004, as the reference computes it):
//   forward  (A x)_k   = conj( E1_k / (4 NX NY) * sum_{l,l'} g_x(l) g_y(l') G[m_k+l, n_k+l'] )
//   adjoint  Re(A^H w r) = E4 / (8 NX NY) * c2r( S ),  S = G_r + conj(G_r(-k)) on ky >= 0
// with G the unnormalised r2c spectrum of the centred, E4-weighted image padded
// to 2NX x 2NY, g(l) = exp(l a - l^2 t3) the Gaussian tap weights, and G_r the
// grid of E1 w conj(r) deposits (r: residual already weighted once by w = 1/sigma).  Every visibility is first reflected to v >= 0
// (u,v -> -u,-v, vis -> conj vis), so all taps lie at ky >= -(MSP-1).
//
// HBM layout (complex = float2/double2):
//   Gf [ky + MSP - 1][kx + NX]   forward spectrum, ky in [-(MSP-1), NY], 2NX columns
//   Hg [ky][kx + NX]              adjoint grid, ky in [0, NY], 2NX columns (kept zero
//                                 between adjoints: the column pass clears it)
//   T  [ky][i]                    forward row-pass output (NY+1 rows of NX)
//   U  [i][ky]                    adjoint column-pass output (NX rows of NY+1)
#pragma once

#include "common.h"
#include "fft_lds.h"

namespace mf {

// Per-visibility parameters in tile-sorted order.
template <typename R>
struct VisParams {
  int2* mxy;      // (mx, my) grid cell after reflection to v >= 0
  cplx<R>* a;     // tap slopes (a_x, a_y)
  R* e1;          // E1
  cplx<R>* vis;   // visibility in the reflected frame
  R* w;           // 1 / sigma
};

struct GridGeom {
  int NX, NY, M, MSP, Lx, Ly;  // Lx = 2NX, Ly = 2NY
  int TX, TY, ntx, nty;        // gridding tiles over Hg
  double t3x, t3y;             // tap curvature (pi/Mr)^2 / tau
  double inv_MM;               // 1 / (4 NX NY)
};

// ---- launchers (all asynchronous on `st`) ----
template <typename R>
void launch_rowfft_fwd(const R* x, const R* e4x, const R* e4y, cplx<R>* T, int NX, int NY, int Ly,
                       int shift, const FFTPlan& P, const cplx<R>* tw, hipStream_t st,
                       const int* skip);
template <typename R>
void launch_colfft_fwd(const cplx<R>* T, cplx<R>* out, int NX, int nky, int Lx, int shift,
                       int mode, int MSP, const FFTPlan& P, const cplx<R>* tw, hipStream_t st,
                       const int* skip);
template <typename R>
void launch_colfft_inv(cplx<R>* in, cplx<R>* U, int NX, int nky, int Lx, int shift,
                       int mode, const FFTPlan& P, const cplx<R>* tw, hipStream_t st,
                       const int* skip);
template <typename R>
void launch_rowfft_c2r(const cplx<R>* U, const R* e4x, const R* e4y, R scale, R* g, int NX,
                       int NY, int Ly, int shift, const FFTPlan& P, const cplx<R>* tw,
                       hipStream_t st, const int* skip);
template <typename R>
void launch_degrid(const GridGeom& g, const VisParams<R>& vp, const cplx<R>* Gf, cplx<R>* rz,
                   float2* rz32, const double2* rz64, double* chi_part, double* lip_part, int mode,
                   int nblk, cplx<R>* model, const int* perm, const int* yneg_sorted,
                   hipStream_t st, const int* skip);
template <typename R>
void launch_grid(const GridGeom& g, const VisParams<R>& vp, const cplx<R>* rz, const int4* chunks,
                 int nchunks, const unsigned* entries, cplx<R>* Hg, hipStream_t st,
                 const int* skip);
constexpr int kGridChunk = 32;  // gridding entries per workgroup

int degrid_blocks(int M, int MSP);

// setup kernels
void launch_vis_keys(const double* u, const double* v, int M, int NX, int NY, int preset,
                     int MSP, int TX, int TY, int ntx, int* keys, int* idx, int* err,
                     hipStream_t st);
void launch_vis_params(const int* perm, const double* u, const double* v, const double* vr,
                       const double* vi, const double* sd, int M, int NX, int NY, int preset,
                       int MSP, VisParams<float> pf, VisParams<double> pd, int* yneg,
                       hipStream_t st);
// Tile-list slots per visibility: the 2 MSP x 2 MSP footprint touches at most
// kx x ky tiles directly and kx x km through its conjugate mirror below ky = 0.
inline int tile_slots(int MSP, int TX, int TY) {
  const int kx = (2 * MSP - 2) / TX + 2, ky = (2 * MSP - 2) / TY + 2, km = MSP / TY + 1;
  return kx * ky + kx * km;
}
void launch_tile_entries(const int2* mxy, int M, int NX, int NY, int MSP, int TX, int TY,
                         int ntx, int nty, int slots, int* keys, unsigned* vals, hipStream_t st);
void launch_tile_ranges(const int* keys_sorted, int n, int ntiles, int* start, int* end,
                        hipStream_t st);

}  // namespace mf
