// toeplitz_body.h -- device code of the persistent Toeplitz-loop MFISTA solver.
//
// For a real image x the data term's normal operator T x = Re A^H W^2 A x is
// a convolution: (T x)(a) = sum_b x(b) p(a - b) with the point-spread function
// p(D) = sum_m w_m^2 cos(u_m . D), D in (-N, N)^2.  Embedded in the circulant
// of the 2NX x 2NY grid it is diagonal in Fourier space, with a real spectrum
// Kh (p is real and even).  So the gradient at z, b - T z with b = Re A^H W^2 V,
// and the objective of a trial,
//   chi^2(x) = |W V|^2 - 2 <x, b> + <x, T x>,
// need no gridding or degridding inside the loop: every trial applies T to its
// step d = xnew - z with an FP64 real-to-complex FFT on the padded grid, the
// product with Kh and the inverse FFT (T xnew = T z + T d, and the momentum
// carries T x and T xnew as it carries x and xnew).  b, p and so Kh come from
// a 24-tap FP64 NUFFT adjoint at set-up (toeplitz.hip); to that transform's
// accuracy the loop is the reference's loop (c/mfista_nufft_lib.c:557-787)
// with its exact operator.
//
// One cooperative launch per solve, L1 and L1+TSV, paired trial rounds.  Per
// round: momentum (first round of an iteration) + gradient + prox + FP64 row
// FFTs of the two steps | column FFT, product with Kh, inverse column FFT |
// c2r rows, T xnew and the trial partials -> control: 3 r grid barriers for
// r rounds (the gridding loop of persistent_body.h: 5 + 3 (r - 1)).  The
// momentum needs no stage of its own: the images live in four buffers whose
// roles rotate (x, the two trials' xnew, a spare), so the first round reads
// x and the accepted xnew, which no workgroup writes in that stage, and forms
// z (also at the neighbours its TSV gradient needs) as it goes.
#pragma once
#include "persistent_body.h"

namespace mf {

// Column pass of the Toeplitz product for ky row `ky` of nseq (1 or 2) row
// transforms T0, T1 (layout T[ky][i], i < NX, as st_rowfft_fwd writes it):
// the rows placed on the Lx-point grid as st_colfft_fwd places them, the
// forward FFT along x, the product with the real kernel row Kh[ky][k], the
// backward FFT and the NX image rows stored transposed, U[i][ky] (as
// st_colfft_inv stores them, for st_rowfft_c2r).
template <typename R>
__device__ inline void st_colconv(int ky, int nseq, const cplx<R>* __restrict__ T0, const cplx<R>* __restrict__ T1,
                                  const R* __restrict__ Kh, cplx<R>* __restrict__ U0, cplx<R>* __restrict__ U1,
                                  int NX, int nky, int Lx, int shift, const FFTPlan& P,
                                  const cplx<R>* __restrict__ tw, cplx<R>* sm) {
  using C = cplx<R>;
  constexpr int kPre = 8;  // kernel values a thread prefetches (Lx <= kPre * blockDim)
  C* b0 = sm;
  C* b1 = sm + nseq * Lx;
  // the kernel row is read before the forward transform, so that its load
  // latency overlaps the transform
  const R* kr = Kh + (size_t)ky * Lx;
  R kv[kPre];
#pragma unroll
  for (int m = 0; m < kPre; m++) {
    const int k = threadIdx.x + m * blockDim.x;
    kv[m] = (k < Lx) ? kr[k] : (R)0;
  }
  for (int t = threadIdx.x; t < nseq * Lx; t += blockDim.x) {
    const int q = t >= Lx, a = t - q * Lx;
    int i = a + shift;
    if (i >= Lx) i -= Lx;
    b0[t] = (i < NX) ? (q ? T1 : T0)[(size_t)ky * NX + i] : mk<R>(0, 0);
  }
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, nseq, P, tw, -1);
  C* oth = (res == b0) ? b1 : b0;
#pragma unroll
  for (int m = 0; m < kPre; m++) {
    const int k = threadIdx.x + m * blockDim.x;
    if (k < Lx) {
      for (int q = 0; q < nseq; q++) {
        const C v = res[q * Lx + k];
        res[q * Lx + k] = mk<R>(v.x * kv[m], v.y * kv[m]);
      }
    }
  }
  __syncthreads();
  res = lds_fft<R>(res, oth, nseq, P, tw, +1);
  for (int t = threadIdx.x; t < nseq * NX; t += blockDim.x) {
    const int q = t >= NX, i = t - q * NX;
    int a = i - shift;
    if (a < 0) a += Lx;
    (q ? U1 : U0)[(size_t)i * nky + ky] = res[q * Lx + a];
  }
  __syncthreads();
}

// st_colconv for one transform, with the two real ky rows packed: task ky = 0
// carries rows 0 and NY (= nky - 1) together, as w = T[0] + i T[NY] (both
// rows of a real image's half spectrum are real).  After the forward FFT the
// two spectra are W(k) = F0(k) + i FN(k) with F0, FN Hermitian, so
//   K0 F0 + i KN FN = W (K0 + KN) / 2 + conj(W(-k)) (K0 - KN) / 2,
// and the backward FFT returns the two rows' products as its real and
// imaginary parts (stored with zero imaginary parts, which the c2r pass
// ignores at ky = 0 and NY).  Tasks ky = 1 .. NY-1 are ordinary columns, so a
// transform takes NY tasks instead of NY + 1.
//
// The Lx = 2N-point transforms are pruned.  The N image rows occupy positions
// a = (i - N/2) mod 2N, one in each residue class mod N, so
//   X(2m) = FFT_N(y)(m),  X(2m+1) = FFT_N(y W_2N^a)(m),  y(a mod N) = x(a),
// two N-point transforms (one LDS stage fewer than one 2N-point transform);
// and the N rows wanted back are e(r) + conj(W_2N^a) o(r), r = a mod N, with
// e and o the N-point backward transforms of the even and odd halves of the
// product (the last radix-2 stage folded into the store).  Ph / twh: the
// N-point plan and twiddles; tw: the 2N-point twiddles.
template <typename R>
__device__ inline void st_colconv1(int ky, const cplx<R>* __restrict__ T, const R* __restrict__ Kh,
                                   cplx<R>* __restrict__ U, int NX, int nky, int Lx, int shift, const FFTPlan& Ph,
                                   const cplx<R>* __restrict__ twh, const cplx<R>* __restrict__ tw, cplx<R>* sm) {
  using C = cplx<R>;
  constexpr int kPre = 4;  // kernel values a thread prefetches (Lx <= kPre * blockDim)
  const int N = Lx / 2, Nh = N / 2;
  const int kyN = nky - 1;
  const bool packed = (ky == 0);
  C* b0 = sm;          // [even | odd] half sequences, N each
  C* b1 = sm + Lx;
  const R* kr = Kh + (size_t)ky * Lx;
  const R* kn = Kh + (size_t)kyN * Lx;
  R kv[kPre], kw[kPre];
#pragma unroll
  for (int m = 0; m < kPre; m++) {
    const int k = threadIdx.x + m * blockDim.x;
    kv[m] = (k < Lx) ? kr[k] : (R)0;
    kw[m] = (packed && k < Lx) ? kn[k] : (R)0;
  }
  // residue r of position a: a = r (r < N/2) or r + N; image row i = (r + N/2) mod N
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    int i = r + Nh;
    if (i >= N) i -= N;
    const int a = (r < Nh) ? r : r + N;
    C v = mk<R>(0, 0);
    if (i < NX) v = packed ? mk<R>(T[i].x, T[(size_t)kyN * NX + i].x) : T[(size_t)ky * NX + i];
    b0[r] = v;
    b0[N + r] = cmul(v, tw[a]);
  }
  (void)shift;  // the placement above is the shift NX / 2 of st_colfft_fwd, with Lx = 2 NX
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, 2, Ph, twh, -1);
  C* oth = (res == b0) ? b1 : b0;
  auto at = [&](int k) -> int { return (k & 1) ? N + (k >> 1) : (k >> 1); };  // X(k) in res
  if (!packed) {
#pragma unroll
    for (int m = 0; m < kPre; m++) {
      const int k = threadIdx.x + m * blockDim.x;
      if (k < Lx) {
        const C v = res[at(k)];
        res[at(k)] = mk<R>(v.x * kv[m], v.y * kv[m]);
      }
    }
  } else {
#pragma unroll
    for (int m = 0; m < kPre; m++) {
      const int k = threadIdx.x + m * blockDim.x;
      if (k < Lx) {
        const C w = res[at(k)], wm = res[at(k == 0 ? 0 : Lx - k)];  // W(k), W(-k)
        const R sp = (R)0.5 * (kv[m] + kw[m]), sn = (R)0.5 * (kv[m] - kw[m]);
        oth[at(k)] = mk<R>(w.x * sp + wm.x * sn, w.y * sp - wm.y * sn);
      }
    }
    C* t = res;  // the product is in oth: transform it, res is the scratch
    res = oth;
    oth = t;
  }
  __syncthreads();
  res = lds_fft<R>(res, oth, 2, Ph, twh, +1);
  for (int i = threadIdx.x; i < NX; i += blockDim.x) {
    int r = i + Nh;
    if (r >= N) r -= N;
    const int a = (r < Nh) ? r : r + N;
    const C o = res[N + r], wa = tw[a];
    const C v = cadd(res[r], cmul(o, mk<R>(wa.x, -wa.y)));
    if (!packed) {
      U[(size_t)i * nky + ky] = v;
    } else {
      U[(size_t)i * nky] = mk<R>(v.x, 0);
      U[(size_t)i * nky + kyN] = mk<R>(v.y, 0);
    }
  }
  __syncthreads();
}

// The row passes of the loop, pruned like st_colconv1's transforms (square
// images, so the NX-point plan Ph / twiddles twh serve the rows too; tw: the
// Ly-point twiddles).  Pixel j of a row sits at b = (j - NY/2) mod Ly, one per
// residue class r = b mod NY, with j = (r + NY/2) mod NY.
__device__ __forceinline__ int toep_pos(int r, int Nh) { return (r < Nh) ? r : r + 2 * Nh; }

// st_rowfft_fwd with nseq = 1 (rows 2 task, 2 task + 1 packed as one complex
// sequence): the even and odd outputs X(2m), X(2m+1) of the Ly-point transform
// are the NY-point transforms of y and y W_Ly^b.  Each pixel is loaded once.
template <typename R, typename Load, bool NTS>
__device__ inline void st_rowfft_fwd_h(int task, int NX, int NY, const FFTPlan& Ph, const cplx<R>* __restrict__ twh,
                                       const cplx<R>* __restrict__ tw, cplx<R>* __restrict__ T, cplx<R>* sm,
                                       Load load) {
  using C = cplx<R>;
  const int N = NY, Nh = N / 2;
  C* b0 = sm;  // [even | odd], N each
  C* b1 = sm + 2 * N;
  const int i0 = 2 * task, i1 = i0 + 1;
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    int j = r + Nh;
    if (j >= N) j -= N;
    R re = 0, im = 0;
    if (i0 < NX) re = load(i0, j);
    if (i1 < NX) im = load(i1, j);
    const C v = mk<R>(re, im);
    b0[r] = v;
    b0[N + r] = cmul(v, tw[toep_pos(r, Nh)]);
  }
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, 2, Ph, twh, -1);
  auto at = [&](int k) -> int { return (k & 1) ? N + (k >> 1) : (k >> 1); };  // X(k) in res
  // half spectra of the two rows, T[k][i], k in [0, NY]
  for (int k = threadIdx.x; k <= N; k += blockDim.x) {
    const C z = res[at(k)], zm = res[at(k == 0 ? 0 : 2 * N - k)];
    const C o0 = mk<R>((R)0.5 * (z.x + zm.x), (R)0.5 * (z.y - zm.y));
    const C o1 = mk<R>((R)0.5 * (z.y + zm.y), (R)-0.5 * (z.x - zm.x));
    cplx<R>* out = T + (size_t)k * NX + i0;
    if constexpr (NTS) {
      nt_store<R>(out, o0);
      if (i1 < NX) nt_store<R>(out + 1, o1);
    } else {
      out[0] = o0;
      if (i1 < NX) out[1] = o1;
    }
  }
  __syncthreads();
}

// st_rowfft_c2r with nseq = 1: the Hermitian-completed spectrum of rows
// 2 task, 2 task + 1 (packed as A + i B) is split into its even and odd
// entries, both go through the NY-point backward transform, and pixel j is
// e(r) + conj(W_Ly^b) o(r).  store(i, j, v) as in st_rowfft_c2r.
template <typename R, typename Store>
__device__ inline void st_rowfft_c2r_h(int task, const cplx<R>* __restrict__ U, int NX, int NY, const FFTPlan& Ph,
                                       const cplx<R>* __restrict__ twh, const cplx<R>* __restrict__ tw, cplx<R>* sm,
                                       Store store) {
  using C = cplx<R>;
  const int N = NY, Nh = N / 2, Ly = 2 * N, nky = N + 1;
  C* b0 = sm;
  C* b1 = sm + Ly;
  const int i0 = 2 * task, i1 = i0 + 1;
  for (int b = threadIdx.x; b < Ly; b += blockDim.x) {
    const int k = (b <= N) ? b : Ly - b;
    C A = (i0 < NX) ? U[(size_t)i0 * nky + k] : mk<R>(0, 0);
    C B = (i1 < NX) ? U[(size_t)i1 * nky + k] : mk<R>(0, 0);
    if (k == 0 || k == N) {
      A.y = 0;
      B.y = 0;
    }
    b0[(b & 1) ? N + (b >> 1) : (b >> 1)] = (b <= N) ? mk<R>(A.x - B.y, A.y + B.x) : mk<R>(A.x + B.y, -A.y + B.x);
  }
  __syncthreads();
  C* res = lds_fft<R>(b0, b1, 2, Ph, twh, +1);
  for (int j = threadIdx.x; j < N; j += blockDim.x) {
    int r = j + Nh;
    if (r >= N) r -= N;
    const C o = res[N + r], wa = tw[toep_pos(r, Nh)];
    const C z = cadd(res[r], cmul(o, mk<R>(wa.x, -wa.y)));
    if (i0 < NX) store(i0, j, z.x);
    if (i1 < NX) store(i1, j, z.y);
  }
  __syncthreads();
}

// LDS bytes of the FP64 twiddle tables: the 2N-point table (square images:
// rows and columns share it) and the pruned transforms' N-point table
__host__ __device__ inline size_t toep_table_bytes(int Lx, int Ly) {
  return (size_t)(Lx + (Ly != Lx ? Ly : 0) + Lx / 2) * 16;
}

// The solve by workgroups lb = 0 .. nb-1 of a cooperative launch.  TAB: the
// twiddle tables are staged in LDS.
template <bool TAB>
__device__ __forceinline__ void toeplitz_body(const PArgs& A, const int lb, const unsigned nb) {
  extern __shared__ double4 smem_d4[];
  __shared__ double red[16 * kPW];
  __shared__ BlockState bs;
  __shared__ double ring[kTD + 1];
  __shared__ BarCtx B;
  const GridGeom& g = A.g;
  const int NX = g.NX, NY = g.NY, n = NX * NY, W = NX, H = NY;
  const int nky = NY + 1;
  const bool tsv = A.lambda_tsv > 0.0;
  unsigned long long epoch = 0;
  double* prow = GP(A.part) + (int)lb;
  const size_t pstr = nb;

  if (!bar_setup(B, GP(A.bar), nb, A.hier_barrier)) return;
  if (threadIdx.x == 0) {
    B.trace = A.trace;
    B.ttag = A.ttag;
    B.tcap = A.tcap;
    B.trow = (int)lb;
    B.tnrow = (int)nb;
  }
  __syncthreads();
  char* lds = reinterpret_cast<char*>(smem_d4);
  const double2 *twxd = GP(A.twxd), *twyd = GP(A.twyd), *twxh = GP(A.twxh);
  if constexpr (TAB) {
    // (the loop takes square images only: the Lx- and Ly-point tables are
    // the same table, staged once; the layout keeps a second one otherwise)
    const bool sq = g.Lx == g.Ly;
    double2* t64x = reinterpret_cast<double2*>(lds);
    double2* t64y = sq ? t64x : t64x + g.Lx;
    double2* t64h = t64x + g.Lx + (sq ? 0 : g.Ly);
    for (int i = threadIdx.x; i < g.Lx; i += blockDim.x) t64x[i] = GP(A.twxd)[i];
    if (!sq)
      for (int i = threadIdx.x; i < g.Ly; i += blockDim.x) t64y[i] = GP(A.twyd)[i];
    for (int i = threadIdx.x; i < g.Lx / 2; i += blockDim.x) t64h[i] = GP(A.twxh)[i];
    twxd = t64x;
    twyd = t64y;
    twxh = t64h;
    lds += (toep_table_bytes(g.Lx, g.Ly) + 255) & ~(size_t)255;
    __syncthreads();
  }
  double2* sm64 = reinterpret_cast<double2*>(lds);
  const double* bv = GP(A.bimg);
  const double* kh = GP(A.kh);
  const float* boxp = GP(A.box);
  double* xs = GP(A.x);
  double* zs = GP(A.z);

  // ---------------- initial cost: z = x0, T x0 -> tv[0] (x) and tv[2] (z)
  {
    double st[2] = {0, 0};
    for (int t = lb; t < A.n_row; t += nb) {
      auto load = [&](int i, int j) -> double {
        const int f = i * NY + j;
        const double xc = xs[f];
        zs[f] = xc;
        GP(A.xb[0])[f] = xc;
        st[0] += fabs(xc);
        const int c = f % W, r = f / W;
        double dh = 0, dv = 0;
        if (c < W - 1) dh = xc - xs[f + 1];
        if (r < H - 1) dv = xc - xs[f + W];
        st[1] += dh * dh + dv * dv;
        return xc;
      };
      st_rowfft_fwd<double>(t, 1, NX, NY, g.Ly, NY / 2, A.Py, twyd, GP(A.T64), sm64, load);
    }
    block_sum<2>(st, red);
    if (threadIdx.x == 0) {
      prow[(kPIl1) * pstr] = st[0];
      prow[(kPItsv) * pstr] = st[1];
    }
  }
  if (!grid_sync(B, epoch, 1)) return;
  for (int t = lb; t < nky; t += nb)
    st_colconv<double>(t, 1, GP(A.T64), GP(A.T64), kh, GP(A.u64a), GP(A.u64a), NX, nky, g.Lx, NX / 2, A.Px, twxd,
                       sm64);
  if (!grid_sync(B, epoch, 2)) return;
  {
    double v[1] = {0};
    double* tx = GP(A.tv[0]);
    double* tz = GP(A.tv[2]);
    for (int t = lb; t < A.n_row; t += nb) {
      auto store = [&](int i, int j, double val) {
        const int f = i * NY + j;
        tx[f] = val;
        tz[f] = val;
        v[0] += xs[f] * (val - 2.0 * bv[f]);
      };
      st_rowfft_c2r<double>(t, 1, GP(A.u64a), NX, NY, g.Ly, NY / 2, A.Py, twyd, sm64, store);
    }
    block_sum<1>(v, red);
    if (threadIdx.x == 0) prow[(kPIchi) * pstr] = v[0];
  }
  if (!grid_sync(B, epoch, 3)) return;
  {
    const int slot[3] = {kPIchi, kPIl1, kPItsv};
    double sum[3];
    part_reduce<3>(GP(A.part), nb, slot, sum, red);
    if (threadIdx.x == 0) {
      double c0 = (A.wv2 + sum[0]) / 2 + A.lambda_l1 * sum[1];
      if (tsv) c0 += A.lambda_tsv * sum[2];
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
      bs.xr[0] = 0;
      bs.xr[1] = 1;
      bs.xr[2] = 2;
      bs.xr[3] = 3;
    }
    __syncthreads();
  }

  // ---------------- the MFISTA loop
  while (true) {
    const int prologue = bs.prologue;
    // momentum of the first round of an iteration: z = a x + b xnew and
    // T z = a T x + b T xnew, with xnew the last accepted-by-the-test trial;
    // x <- xnew when its objective decreased.  The roles read here are the
    // ones before the rotation below.
    const bool mom = prologue && bs.mom_pending;
    const double ma = bs.mom_a, mb = bs.mom_b;
    const int ka = bs.kacc;
    const double* xo = GP(A.xb[bs.xr[0]]);            // x
    const double* xq = GP(A.xb[bs.xr[ka ? 2 : 1]]);   // xnew
    const double* tx = GP(A.tv[bs.ix]);
    const double* tq = GP(A.tv[ka ? bs.in1 : bs.in]);
    __syncthreads();  // every thread has read the roles
    if (mom && threadIdx.x == 0) {
      const int rx = bs.xr[0], rq = bs.xr[ka ? 2 : 1], ro = bs.xr[ka ? 1 : 2], rs = bs.xr[3];
      if (bs.accept) {  // x <- xnew (buffer rq); the old x is not written before the next momentum
        bs.xr[0] = rq;
        bs.xr[3] = rx;
        int& rn = ka ? bs.in1 : bs.in;  // T x <- T xnew: swap roles
        const int tmp = bs.ix;
        bs.ix = rn;
        rn = tmp;
      } else {
        bs.xr[3] = rq;
      }
      bs.xr[1] = ro;  // the trials write the two buffers nobody reads in this stage
      bs.xr[2] = rs;
    }
    __syncthreads();
    // T1: gradient at z, the two trials' prox images and steps, FP64 row FFTs
    // of the steps (c and c * eta side by side)
    const double c = bs.c;
    const double etaf = (bs.iter == 0) ? A.eta_first : kEta;
    const double c1 = c * etaf;  // the value the control step gives c after a failed trial
    const double ci0 = 1.0 / c, e0 = A.lambda_l1 / c, ci1 = 1.0 / c1, e1 = A.lambda_l1 / c1;
    const double lt = A.lambda_tsv;
    double* tz = GP(A.tv[bs.iz]);
    double* xa_out = GP(A.xb[bs.xr[1]]);
    double* xb_out = GP(A.xb[bs.xr[2]]);
    {
      // task 2 q + k: row pair q (rows 2q, 2q+1), trial k -- two workgroups
      // per row pair, each forming the gradient and taking one trial's prox
      // and FFT.  z of rows 2q-1 .. 2q+2 (the pair and its TSV neighbours),
      // formed from x and xnew in the momentum round, is staged in the upper
      // half of the row pass's LDS scratch (unused until its first FFT stage:
      // 4 NY doubles = Ly complex doubles)
      double v[4] = {0, 0, 0, 0};
      double* zl = reinterpret_cast<double*>(sm64 + g.Ly);
      for (int t = lb; t < 2 * A.n_row; t += nb) {
        const int k = t & 1, rt = t >> 1;
        const int r0 = 2 * rt - 1;
        const double cik = k ? ci1 : ci0, ek = k ? e1 : e0;
        double* xk_out = k ? xb_out : xa_out;
        // T z and b at this thread's pixels of the row transform's load (r =
        // threadIdx.x + m blockDim, pixel j = (r + NY/2) mod NY, rows 2 rt + q),
        // read before the staging barrier so that their latency overlaps the
        // staging loads (NY <= 2 blockDim)
        double pt[2][2] = {{0, 0}, {0, 0}}, pb[2][2] = {{0, 0}, {0, 0}};
#pragma unroll
        for (int m = 0; m < 2; m++) {
          const int r = threadIdx.x + m * blockDim.x;
          if (r < NY) {
            int j = r + NY / 2;
            if (j >= NY) j -= NY;
#pragma unroll
            for (int q = 0; q < 2; q++) {
              const int i = 2 * rt + q;
              if (i < NX) {
                const int f = i * NY + j;
                pt[m][q] = mom ? ma * tx[f] + mb * tq[f] : tz[f];
                pb[m][q] = bv[f];
              }
            }
          }
        }
        for (int q = threadIdx.x; q < 4 * NY; q += blockDim.x) {
          const int r = r0 + q / NY, j = q - (q / NY) * NY;
          double zv = 0.0;
          if (r >= 0 && r < NX) {
            const int f = r * NY + j;
            zv = mom ? ma * xo[f] + mb * xq[f] : zs[f];
          }
          zl[q] = zv;
        }
        __syncthreads();
        double d2 = 0.0, l1 = 0.0;
        auto load = [&](int i, int j) -> double {
          const int f = i * NY + j;
          const int l = (i - r0) * NY + j;  // (i, j) in the staged rows
          const double zf = zl[l];
          int rr = j + NY / 2;  // the load's r for pixel j
          if (rr >= NY) rr -= NY;
          const bool m1 = rr >= (int)blockDim.x, q1 = i != 2 * rt;
          const double tzf = m1 ? (q1 ? pt[1][1] : pt[1][0]) : (q1 ? pt[0][1] : pt[0][0]);
          const double bf = m1 ? (q1 ? pb[1][1] : pb[1][0]) : (q1 ? pb[0][1] : pb[0][0]);
          if (mom && !k) {  // the trial-0 workgroup of the pair stores z and T z
            zs[f] = zf;
            tz[f] = tzf;
          }
          double gf = bf - tzf;
          if (tsv) {  // - lambda_tsv grad TSV(z); square images: column j, row i
            double d = 0.0;
            if (j < W - 1) d += zf - zl[l + 1];
            if (j > 0) d += zf - zl[l - 1];
            if (i < H - 1) d += zf - zl[l + NY];
            if (i > 0) d += zf - zl[l - NY];
            gf -= lt * (2.0 * d);
          }
          const bool boxed = boxp && boxp[f] == 0.0f;
          const double xv = prox_p<double>(gf * cik + zf, ek, A.nonneg, boxed);
          xk_out[f] = xv;
          const double dv = xv - zf;
          d2 += dv * dv;
          l1 += fabs(xv);
          return dv;
        };
        // (non-temporal stores of the row transforms: measured T1 barrier
        // tail 5.0 -> 3.6 us, the release has fewer dirty lines to write back)
        st_rowfft_fwd_h<double, decltype(load), true>(rt, NX, NY, A.Pxh, twxh, twyd, k ? GP(A.T64b) : GP(A.T64),
                                                      sm64, load);
        if (k) {  // (no runtime index into v: it stays in registers)
          v[2] += d2;
          v[3] += l1;
        } else {
          v[0] += d2;
          v[1] += l1;
        }
      }
      block_sum<4>(v, red);
      if (threadIdx.x == 0) {
        prow[(kPD2) * pstr] = v[0];
        prow[(kPL1) * pstr] = v[1];
        prow[(kPTrial + kPD2) * pstr] = v[2];
        prow[(kPTrial + kPL1) * pstr] = v[3];
      }
    }
    if (!grid_sync(B, epoch, 8)) return;
    // T2: column FFT, product with Kh, inverse column FFT; task 2 q + k is
    // ky row q of trial k (row 0 carrying the Nyquist row too)
    for (int t = lb; t < 2 * NY; t += nb) {
      const int k = t & 1;
      st_colconv1<double>(t >> 1, k ? GP(A.T64b) : GP(A.T64), kh, k ? GP(A.u64b) : GP(A.u64a), NX, nky, g.Lx,
                          NX / 2, A.Pxh, twxh, twxd, sm64);
    }
    if (!grid_sync(B, epoch, 9)) return;
    // T3: c2r rows of both trials: T d, T xnew = T z + T d, chi^2(xnew) - |W V|^2
    // and |W A d|^2 = <d, T d>; task 2 q + k is row task q of trial k.  The
    // regulariser partials of xnew and d: on the workgroups without a row
    // task when at least half as many of them as pixels / 256 exist, else on
    // every workgroup after its row tasks
    {
      const int ntask = 2 * A.n_row;
      const int nreg = (int)nb - ntask;
      const bool split = nreg > 0 && 2 * (long)nreg * blockDim.x >= n;
      const bool do_reg = !split || (int)lb >= ntask;
      const int rg0 = split ? (int)lb - ntask : (int)lb;
      const int rgn = split ? nreg : (int)nb;
      double r[4] = {0, 0, 0, 0};
      if (do_reg) {  // (uniform per workgroup: the reduction's barriers are safe)
        for (int f = rg0 * blockDim.x + threadIdx.x; f < n; f += rgn * blockDim.x) {
          reg_terms(xa_out, zs, f, W, H, false, r);
          reg_terms(xb_out, zs, f, W, H, false, r + 2);
        }
        block_sum<4>(r, red);
      }
      if (threadIdx.x == 0) {
        prow[(kPReg) * pstr] = r[0];
        prow[(kPTsvd) * pstr] = r[1];
        prow[(kPTrial + kPReg) * pstr] = r[2];
        prow[(kPTrial + kPTsvd) * pstr] = r[3];
      }
      double v[4] = {0, 0, 0, 0};
      for (int t = lb; t < 2 * A.n_row; t += nb) {
        const int k = t & 1, rt = t >> 1;
        const double* xk = k ? xb_out : xa_out;
        double* tk = GP(A.tv[k ? bs.in1 : bs.in]);
        double chi = 0.0, lip = 0.0;
        auto store = [&](int i, int j, double td) {
          const int f = i * NY + j;
          const double xv = xk[f];
          const double tn = tz[f] + td;
          tk[f] = tn;
          lip += (xv - zs[f]) * td;
          chi += xv * (tn - 2.0 * bv[f]);
        };
        st_rowfft_c2r_h<double>(rt, k ? GP(A.u64b) : GP(A.u64a), NX, NY, A.Pxh, twxh, twyd, sm64, store);
        if (k) {  // (no runtime index into v: it stays in registers)
          v[2] += chi;
          v[3] += lip;
        } else {
          v[0] += chi;
          v[1] += lip;
        }
      }
      block_sum<4>(v, red);
      if (threadIdx.x == 0) {
        prow[(kPChi) * pstr] = v[0];
        prow[(kPLip) * pstr] = v[1];
        prow[(kPTrial + kPChi) * pstr] = v[2];
        prow[(kPTrial + kPLip) * pstr] = v[3];
      }
    }
    if (!grid_sync(B, epoch, 10)) return;
    // ---------------- control (every workgroup, identical arithmetic)
    double sum[2 * kPTrial];
    {
      const int sl[12] = {kPChi, kPLip, kPD2, kPL1, kPReg, kPTsvd,
                          kPTrial + kPChi, kPTrial + kPLip, kPTrial + kPD2, kPTrial + kPL1,
                          kPTrial + kPReg, kPTrial + kPTsvd};
      double s1[12];
      part_reduce<12>(GP(A.part), nb, sl, s1, red);
#pragma unroll
      for (int k = 0; k < 2; k++) {
#pragma unroll
        for (int q = 0; q < 6; q++) sum[k * kPTrial + q] = s1[k * 6 + q];
        sum[k * kPTrial + kPTx] = 0.0;
        sum[k * kPTrial + kPTd] = 0.0;
      }
    }
    if (threadIdx.x == 0) round_decide(A, bs, ring, sum, true, false, tsv, false, (int)lb, prologue, A.wv2);
    __syncthreads();
    if (bs.stop) break;
  }

  // the final image (x, or the accepted image of the last iteration: the
  // reference's x = xnew) into the output buffer
  {
    const int rf = (bs.mom_pending && bs.accept) ? bs.xr[bs.kacc ? 2 : 1] : bs.xr[0];
    const double* xf = GP(A.xb[rf]);
    for (int f = lb * blockDim.x + threadIdx.x; f < n; f += nb * blockDim.x) xs[f] = xf[f];
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

template <bool TAB>
__global__ void __launch_bounds__(MF_BLOCK) k_mfista_toeplitz(const PArgs* __restrict__ A) {
  toeplitz_body<TAB>(*A, blockIdx.x, gridDim.x);
}

// independent solves in one cooperative launch: blk[b] = (problem, workgroup
// index within the problem) for workgroup b
template <bool TAB>
__global__ void __launch_bounds__(MF_BLOCK) k_mfista_toeplitz_batch(const PArgs* __restrict__ args,
                                                                    const int2* __restrict__ blk) {
  const int2 pb = blk[blockIdx.x];
  if (pb.x < 0) return;  // a workgroup slot no problem was dealt
  const PArgs& A = args[pb.x];
  toeplitz_body<TAB>(A, pb.y, (unsigned)A.nblk);
}

}  // namespace mf
