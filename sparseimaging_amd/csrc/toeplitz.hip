// toeplitz.hip -- set-up of the Toeplitz loop (toeplitz_body.h): the real
// spectrum Kh of the data term's normal operator on the 2NX x 2NY grid and
// b = Re A^H W^2 V.
//
// Both come from two adjoints of one 24-tap FP64 NUFFT on a 2L x 2L image,
// L = max(NX, NY), whose pixel (I, J) sits at offset (I - L, J - L):
//   p(D) = Re A_2L^H W w   (the point-spread function sum_m w_m^2 cos(u_m . D)),
//   b    = Re A_2L^H W (w V) cropped to the NX x NY image.
// The 24-tap transform reproduces the exact sums to about 1e-11 (relative),
// so Kh and b describe the exact operator; the 12-tap preset's own error is
// about 1e-6.  p is embedded in the circulant of the Lx x Ly grid (the row
// D = NX and the column D = NY, offsets no pixel pair has, set to 0, so the
// embedding stays even) and transformed with the FP64 stream kernels.
#include <algorithm>

#include "driver.h"
#include "nufft.h"

namespace mf {

int upload_twiddles_pub(const FFTPlan& P, DevMem& tf, DevMem& td, hipStream_t st);  // solver.hip

// the two right-hand sides in the set-up transform's visibility frame
__global__ void k_toep_rhs(VisParams<double> vp, int M, double2* __restrict__ rb, double2* __restrict__ rp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= M) return;
  const double w = vp.w[s];
  const double2 V = vp.vis[s];
  rb[s] = make_double2(w * V.x, w * V.y);
  rp[s] = make_double2(w, 0.0);
}

// c[a][b'] = p(Da, Db') on the Lx x Ly grid (D = a for a < NX, a - Lx above;
// likewise for b'), and b cropped from the 2L image: pixel (i, j) of the
// NX x NY image sits at offset (i - NX/2, j - cy)
__global__ void k_toep_embed(const double* __restrict__ imgp, const double* __restrict__ imgb, int L, int NX, int NY,
                             int cy, double* __restrict__ c, double* __restrict__ b) {
  const int Lx = 2 * NX, Ly = 2 * NY, L2 = 2 * L;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < Lx * Ly) {
    const int a = t / Ly, bb = t - a * Ly;
    double v = 0.0;
    if (a != NX && bb != NY) {
      const int dx = a < NX ? a : a - Lx, dy = bb < NY ? bb : bb - Ly;
      v = imgp[(size_t)(dx + L) * L2 + (dy + L)];
    }
    c[t] = v;
  }
  if (t < NX * NY) {
    const int i = t / NY, j = t - i * NY;
    b[t] = imgb[(size_t)(i - NX / 2 + L) * L2 + (j - cy + L)];
  }
}

// Kh = Re F / (Lx Ly): the circulant's spectrum (real: the embedding is even),
// scaled for the unnormalised forward and backward transforms of the loop
__global__ void k_toep_khat(const double2* __restrict__ F, int n, double scale, double* __restrict__ kh) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) kh[t] = F[t].x * scale;
}

int toeplitz_setup(Context& C, const double* d_u, const double* d_v, const double* d_vr, const double* d_vi,
                   const double* d_sd, int M, int NX, int NY, int preset, double wv2, ToepOut& out) {
  ToepWork& T = C.tw;
  const int L = std::max(NX, NY), L2 = 2 * L;
  if (L2 > 1024 || NX != NY || NX % 2 || !C.nu.ready || C.nu.NX != NX || C.nu.NY != NY) return MFISTA_E_SIZE;
  hipStream_t st = C.stream;
  const size_t img2 = (size_t)L2 * L2 * 8;
  if (T.rb.ensure((size_t)M * 16) || T.rp.ensure((size_t)M * 16) || T.imgp.ensure(img2) || T.imgb.ensure(img2))
    return MFISTA_E_HIP;
  // the set-up transform stands in for C.nu during its calls
  std::swap(C.nu, T.nu);
  int rc = nufft_setup(C, MFISTA_PRESET_C24, d_u, d_v, d_vr, d_vi, d_sd, M, L2, L2);
  if (!rc) rc = nufft_ensure64(C);
  if (!rc) {
    hipLaunchKernelGGL(k_toep_rhs, dim3((M + 255) / 256), dim3(256), 0, st, C.nu.pd(), M, (double2*)T.rb.p,
                       (double2*)T.rp.p);
    nufft_adj<double>(C, (const double2*)T.rp.p, (double*)T.imgp.p, nullptr);
    nufft_adj<double>(C, (const double2*)T.rb.p, (double*)T.imgb.p, nullptr);
    if (hipGetLastError() != hipSuccess) rc = MFISTA_E_HIP;
  }
  std::swap(C.nu, T.nu);
  if (rc) return rc;
  // circulant embedding, its 2-D transform (row pass along y, column pass
  // along x, plain layout F[ky][kx]) and Kh
  NufftState& N = C.nu;
  const int Lx = 2 * NX, Ly = 2 * NY, nky = NY + 1;
  const size_t ngrid = (size_t)Lx * Ly, nspec = (size_t)nky * Lx;
  if (T.circ.ensure(ngrid * 8) || T.tbuf.ensure(nspec * 16) || T.fbuf.ensure(nspec * 16) ||
      out.kh.ensure(nspec * 8) || out.bimg.ensure((size_t)NX * NY * 8))
    return MFISTA_E_HIP;
  const int cy = (preset == MFISTA_PRESET_C24) ? NX / 2 : NY / 2;  // the C library's centring
  hipLaunchKernelGGL(k_toep_embed, dim3((unsigned)((ngrid + 255) / 256)), dim3(256), 0, st,
                     (const double*)T.imgp.p, (const double*)T.imgb.p, L, NX, NY, cy, (double*)T.circ.p,
                     (double*)out.bimg.p);
  launch_rowfft_fwd<double>((const double*)T.circ.p, nullptr, nullptr, (double2*)T.tbuf.p, Lx, Ly, Ly, 0, N.Py,
                            (const double2*)N.twyd.p, st, nullptr);
  launch_colfft_fwd<double>((const double2*)T.tbuf.p, (double2*)T.fbuf.p, Lx, nky, Lx, 0, 1, N.MSP, N.Px,
                            (const double2*)N.twxd.p, st, nullptr);
  hipLaunchKernelGGL(k_toep_khat, dim3((unsigned)((nspec + 255) / 256)), dim3(256), 0, st,
                     (const double2*)T.fbuf.p, (int)nspec, 1.0 / ((double)Lx * Ly), (double*)out.kh.p);
  MF_CHECK(hipGetLastError());
  out.wv2 = wv2;
  // the NX-point plan and twiddles of the loop's pruned column transforms
  if (!make_fft_plan(NX, &out.Pxh)) return MFISTA_E_SIZE;
  if (upload_twiddles_pub(out.Pxh, out.twxh32, out.twxh64, st)) return MFISTA_E_HIP;
  return 0;
}

}  // namespace mf
