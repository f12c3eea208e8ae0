// driver.hip -- the MFISTA outer loop for every mode, the FGP TV prox kernels,
// and the FFT- and DFT-mode operators with their C ABI entry points.
//
// The loop is the reference's (c/mfista_nufft_lib.c:557-787 for L1+TSV,
// :308-553 for L1+TV; the FFT and DFT libraries repeat it with their own
// operators, c/mfista_fft_lib.c:179-609 and c/mfista_dft_lib.c:70-436), run
// as the device state machine of mfista_kernels.hip.  Each enqueued step is
// one backtracking trial; the host polls the `done` flag every 16 steps
// through pinned memory, two polls in flight, so the GPU never idles on it.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "driver.h"

namespace mf {

// ---------------------------------------------------------------------------
// FGP TV prox (Beck & Teboulle), TV view of the flat image: W = NX columns,
// H = NY rows, pixel f = j W + i; duals p[(W-1) j + i] (i < W-1) and
// q[W j + i] (j < H-1) as in c/mfista_TV_lib.c:48-75.

template <typename R>
__device__ inline R fgp_P(R v, R thr, int nonneg, bool boxed) {
  if (boxed) return (R)0;
  if (nonneg) return v < (R)0 ? (R)0 : v;  // P_positive_box, c/mfista_TV_lib.c:84-93
  if (v >= thr) return v - thr;            // soft_threshold_box, c/mfista_tools.c:95-112
  if (v <= -thr) return v + thr;
  return (R)0;
}

template <typename R>
__device__ inline R lpq_at(const R* p, const R* q, int i, int j, int W, int H) {
  R v = 0;
  if (i < W - 1) v += p[(W - 1) * j + i];
  if (i > 0) v -= p[(W - 1) * j + i - 1];
  if (j < H - 1) v += q[W * j + i];
  if (j > 0) v -= q[W * (j - 1) + i];
  return v;
}

// b = z + g/c   (nonneg: b = -l1/c + g/c + z, c/mfista_nufft_lib.c:424-440)
template <typename R>
__global__ void k_tv_pre(const R* __restrict__ z, const R* __restrict__ g, R* __restrict__ b, int n,
                         const DevState* __restrict__ S, const int* skip) {
  if (skip && *skip) return;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const double c = S->c;
  const R cinv = (R)(1.0 / c);
  if (S->nonneg == 1)
    b[f] = (R)(-S->lambda_l1 / c) + cinv * g[f] + z[f];
  else
    b[f] = g[f] * cinv + z[f];
}

// one FGP iteration (c/mfista_TV_lib.c:208-244): x = P(b - lam L(r,s));
// (r,s) += Lx(x)/(8 lam); (np,nq) = P_pqiso(r,s); (r,s) = np c1 + c2 (p,q); (p,q) = (np,nq)
template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_fgp_iter(const R* __restrict__ b, const float* __restrict__ box, const R* __restrict__ rin,
           const R* __restrict__ sin_, R* __restrict__ rout, R* __restrict__ sout, R* __restrict__ p,
           R* __restrict__ q, int W, int H, const DevState* __restrict__ S, R c1, R c2, int first,
           const int* skip) {
  if (skip && *skip) return;
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= W * H) return;
  const int i = f % W, j = f / W;
  const double c = S->c;
  const R lam = (R)(S->lambda_tv / c);
  const R thr = (R)(S->lambda_l1 / c);
  const int nn = (S->nonneg == 1);
  auto xat = [&](int ii, int jj) -> R {
    const int ff = jj * W + ii;
    const R L = first ? (R)0 : lpq_at<R>(rin, sin_, ii, jj, W, H);
    return fgp_P<R>(b[ff] - lam * L, thr, nn, box && box[ff] == 0.0f);
  };
  const R xc = xat(i, j);
  const R a3 = (R)1 / ((R)8 * lam);
  R rr = 0, ss = 0;
  const bool hp = i < W - 1, hq = j < H - 1;
  if (hp) rr = (first ? (R)0 : rin[(W - 1) * j + i]) + (xc - xat(i + 1, j)) * a3;
  if (hq) ss = (first ? (R)0 : sin_[W * j + i]) + (xc - xat(i, j + 1)) * a3;
  R np_ = 0, nq_ = 0;
  if (hp && hq) {
    R t = sqrt(rr * rr + ss * ss);
    if (t < (R)1) t = (R)1;
    np_ = rr / t;
    nq_ = ss / t;
  } else if (hp) {
    R t = fabs(rr);
    if (t < (R)1) t = (R)1;
    np_ = rr / t;
  } else if (hq) {
    R t = fabs(ss);
    if (t < (R)1) t = (R)1;
    nq_ = ss / t;
  }
  if (hp) {
    const int k = (W - 1) * j + i;
    const R pold = first ? (R)0 : p[k];
    rout[k] = np_ * c1 + c2 * pold;
    p[k] = np_;
  }
  if (hq) {
    const int k = W * j + i;
    const R qold = first ? (R)0 : q[k];
    sout[k] = nq_ * c1 + c2 * qold;
    q[k] = nq_;
  }
}

// x = P(b - lam L(p,q)) (c/mfista_TV_lib.c:246-251) and the trial partials:
// <xnew-z, g>, |xnew-z|^2, |xnew|_1, TV(xnew) (c/mfista_TV_lib.c:30-46)
template <typename R>
__global__ void __launch_bounds__(MF_BLOCK)
k_fgp_final(const R* __restrict__ b, const float* __restrict__ box, const R* __restrict__ p,
            const R* __restrict__ q, const R* __restrict__ z, const R* __restrict__ g,
            R* __restrict__ xnew, R* __restrict__ dout, int W, int H,
            const DevState* __restrict__ S, double* part, int have_dual, const int* skip) {
  if (skip && *skip) return;
  __shared__ double red[16 * kProxW];
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  const double c = S->c;
  const R lam = (R)(S->lambda_tv / c);
  const R thr = (R)(S->lambda_l1 / c);
  const int nn = (S->nonneg == 1);
  double v[kProxW] = {0, 0, 0, 0, 0, 0};
  if (f < W * H) {
    const int i = f % W, j = f / W;
    auto xat = [&](int ii, int jj) -> R {
      const int ff = jj * W + ii;
      const R L = have_dual ? lpq_at<R>(p, q, ii, jj, W, H) : (R)0;
      return fgp_P<R>(b[ff] - lam * L, thr, nn, box && box[ff] == 0.0f);
    };
    const R xc = xat(i, j);
    xnew[f] = xc;
    const R d = xc - z[f];
    if (dout) dout[f] = d;
    v[0] = (double)g[f] * (double)d;
    v[1] = (double)d * (double)d;
    v[2] = fabs((double)xc);
    const bool hp = i < W - 1, hq = j < H - 1;
    const double dh = hp ? (double)xc - (double)xat(i + 1, j) : 0.0;
    const double dv = hq ? (double)xc - (double)xat(i, j + 1) : 0.0;
    if (hp && hq)
      v[3] = sqrt(dh * dh + dv * dv);
    else if (hp)
      v[3] = fabs(dh);
    else if (hq)
      v[3] = fabs(dv);
  }
  block_sum<kProxW>(v, red);
  if (threadIdx.x == 0)
    for (int k = 0; k < kProxW; k++) part[kProxW * blockIdx.x + k] = v[k];
}

template <typename R>
void launch_tv_pre(const R* z, const R* g, R* b, int n, const DevState* S, hipStream_t st,
                   const int* skip) {
  hipLaunchKernelGGL(k_tv_pre<R>, dim3(elem_blocks(n)), dim3(MF_BLOCK), 0, st, z, g, b, n, S, skip);
}
template <typename R>
void launch_fgp_iter(const R* b, const float* box, const R* rin, const R* sin, R* rout, R* sout,
                     R* p, R* q, int W, int H, const DevState* S, double c1, double c2, int first,
                     hipStream_t st, const int* skip) {
  hipLaunchKernelGGL(k_fgp_iter<R>, dim3(elem_blocks(W * H)), dim3(MF_BLOCK), 0, st, b, box, rin, sin,
                     rout, sout, p, q, W, H, S, (R)c1, (R)c2, first, skip);
}
template <typename R>
void launch_fgp_final(const R* b, const float* box, const R* p, const R* q, const R* z,
                      const R* g, R* xnew, R* d, int W, int H, const DevState* S, double* part,
                      int have_dual, hipStream_t st, const int* skip) {
  hipLaunchKernelGGL(k_fgp_final<R>, dim3(elem_blocks(W * H)), dim3(MF_BLOCK), 0, st, b, box, p, q, z,
                     g, xnew, d, W, H, S, part, have_dual, skip);
}

// ---------------------------------------------------------------------------
template <typename R>
int mfista_run(Context& C, Operator<R>& op, const RunParams& P, RunStats* out) {
  hipStream_t st = C.stream;
  const int NX = P.NX, NY = P.NY, n = NX * NY;
  const int W = NX, H = NY;  // x[NX*j+i] view of c/mfista_tools.c:145-186
  PartialLayout L;
  L.n_degrid = op.npart();
  L.n_elem = elem_blocks(n);
  L.chi_z = 0;
  L.chi_x = L.n_degrid;
  L.lip = 2 * L.n_degrid;
  L.tsv_z = 3 * L.n_degrid;
  L.prox = L.tsv_z + L.n_elem;
  L.total = L.prox + kProxW * L.n_elem;
  L.n_chiz = L.n_degrid;
  const bool delta = op.has_delta();
  const size_t rb = (size_t)n * sizeof(R);
  if (C.x.ensure(rb) || C.z.ensure(rb) || C.xnew.ensure(rb) || C.g.ensure(rb) || C.dstep.ensure(rb) ||
      C.part.ensure((size_t)L.total * 8) || C.stat.ensure((size_t)4 * L.n_elem * 8) ||
      C.state.ensure(sizeof(DevState)) ||
      C.cost.ensure((size_t)(P.maxiter > 0 ? P.maxiter : 1) * 8))
    return MFISTA_E_HIP;
  if (P.tv_mode && C.fgp.ensure(7 * rb)) return MFISTA_E_HIP;
  R* x = (R*)C.x.p;
  R* z = (R*)C.z.p;
  R* xnew = (R*)C.xnew.p;
  R* g = (R*)C.g.p;
  R* dstep = (R*)C.dstep.p;
  double* part = (double*)C.part.p;
  DevState* S = (DevState*)C.state.p;
  R* fb = (R*)C.fgp.p;  // b, rA, sA, rB, sB, p, q
  R *tb = fb, *rA = fb + n, *sA = fb + 2 * (size_t)n, *rB = fb + 3 * (size_t)n,
    *sB = fb + 4 * (size_t)n, *pp = fb + 5 * (size_t)n, *qq = fb + 6 * (size_t)n;
  launch_cast<double, R>(P.xinit, x, n, st);
  MF_CHECK(hipMemcpyAsync(z, x, rb, hipMemcpyDeviceToDevice, st));
  DevState h;
  memset(&h, 0, sizeof(h));
  h.c = P.cinit;
  h.lambda_l1 = P.lambda_l1;
  h.lambda_tsv = P.tv_mode ? 0.0 : P.lambda_tsv;
  h.lambda_tv = P.tv_mode ? P.lambda_tv : 0.0;
  h.eps_abs = P.eps_abs;
  h.rel_tol = P.rel_tol;
  h.eta_first = P.eta_first;
  h.maxiter = P.maxiter;
  h.nonneg = P.nonneg;
  h.tv_mode = P.tv_mode;
  h.lip_mode = delta ? 1 : 0;
  MF_CHECK(hipMemcpyAsync(S, &h, sizeof(h), hipMemcpyHostToDevice, st));
  op.fwd(x, false, part + L.chi_x, nullptr);
  launch_stats<R>(x, W, H, (double*)C.stat.p, st);
  launch_init_control(S, part, L, (const double*)C.stat.p, L.n_elem, st);
  // FGP momentum coefficients: t_{k+1} = (1 + sqrt(1 + 4 t_k^2)) / 2, t_0 = 1
  std::vector<double> c1(P.fgp_iter > 0 ? P.fgp_iter : 1), c2(c1.size());
  {
    double t = 1;
    for (int it = 0; it < P.fgp_iter; it++) {
      const double tn = (1 + sqrt(1 + 4 * t * t)) / 2;
      c1[it] = 1 + (t - 1) / tn;
      c2[it] = -(t - 1) / tn;
      t = tn;
    }
  }
  const int* gp = &S->gate_prologue;
  const int* gt = &S->gate_trial;
  const float* box = P.box_flag ? P.box : nullptr;
  const int K = 16;
  int* hflag = C.hflags;
  int pending = -1;
  const long cap = (long)P.maxiter * 64 + 100000;
  long step = 0;
  for (;; step++) {
    op.fwd(z, true, part + L.chi_z, gp);
    op.adj(g, gp);
    if (!P.tv_mode && P.lambda_tsv > 0)
      launch_tsv_grad<R>(z, g, W, H, P.lambda_tsv, part + L.tsv_z, st, gp);
    if (!P.tv_mode) {
      launch_prox<R>(z, g, xnew, dstep, box, W, H, S, part + L.prox, st, gt);
    } else {
      launch_tv_pre<R>(z, g, tb, n, S, st, gt);
      R *ri = rA, *si = sA, *ro = rB, *so = sB;
      for (int it = 0; it < P.fgp_iter; it++) {
        launch_fgp_iter<R>(tb, box, ri, si, ro, so, pp, qq, W, H, S, c1[it], c2[it], it == 0, st, gt);
        R* t1 = ri;
        ri = ro;
        ro = t1;
        t1 = si;
        si = so;
        so = t1;
      }
      launch_fgp_final<R>(tb, box, pp, qq, z, g, xnew, dstep, W, H, S, part + L.prox, P.fgp_iter > 0,
                          st, gt);
    }
    if (delta)
      op.fwd_delta(dstep, part + L.chi_x, part + L.lip, gt);
    else
      op.fwd(xnew, false, part + L.chi_x, gt);
    launch_control(S, (double*)C.cost.p, part, L, st);
    launch_momentum<R>(x, xnew, z, n, S, st);
    if ((step + 1) % K == 0) {
      const int slot = (int)((step / K) & 1);
      MF_CHECK(hipMemcpyAsync(&hflag[slot], &S->done, 4, hipMemcpyDeviceToHost, st));
      MF_CHECK(hipEventRecord(C.ev[slot], st));
      if (pending >= 0) {
        MF_CHECK(hipEventSynchronize(C.ev[pending]));
        if (hflag[pending]) break;
      }
      pending = slot;
    }
    if (step > cap) {
      fprintf(stderr, "[mfista_amd] solver did not terminate\n");
      return MFISTA_E_ARG;
    }
  }
  MF_CHECK(hipStreamSynchronize(st));
  MF_CHECK(hipGetLastError());
  MF_CHECK(hipMemcpy(&h, S, sizeof(h), hipMemcpyDeviceToHost));
  out->ITER = h.ITER;
  out->c = h.c;
  out->max_trials = h.max_trials;
  out->total_trials = h.total_trials;
  out->status = h.status;
  out->steps = step + 1;
  launch_cast<R, double>(x, P.xout, n, st);
  MF_CHECK(hipStreamSynchronize(st));
  return 0;
}

template int mfista_run<float>(Context&, Operator<float>&, const RunParams&, RunStats*);
template int mfista_run<double>(Context&, Operator<double>&, const RunParams&, RunStats*);

// ---------------------------------------------------------------------------
// FFT mode (c/mfista_fft_lib.c): (A x) = fft2(x)/sqrt(NN) on the grid cells,
// F = 1/4 sum over the full grid of |mask (y - A x)|^2 rebuilt from the
// ky <= NY/2 half plane, -dF/dx = c2r(mask r / (2 sqrt(NN))).  FP64.
// Half-plane arrays use the layout [ky][kx].

__global__ void k_fft_resid(const double2* __restrict__ X, const double2* __restrict__ yf,
                            const double* __restrict__ mask, double2* __restrict__ rg, int NX,
                            int nky, int NY, double inv_sqrtNN, double* part, const int* skip) {
  if (skip && *skip) return;
  __shared__ double red[16];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {0.0};
  if (t < NX * nky) {
    const int ky = t / NX;
    const double mk = mask[t];
    const double2 x = X[t];
    const double2 y = yf[t];
    const double rr = y.x - mk * x.x * inv_sqrtNN, ri = y.y - mk * x.y * inv_sqrtNN;
    const double w = (ky == 0 || 2 * ky == NY) ? 1.0 : 2.0;
    v[0] = 0.5 * w * (rr * rr + ri * ri);
    if (rg) {
      const double s = mk * 0.5 * inv_sqrtNN;
      rg[t] = make_double2(rr * s, ri * s);
    }
  }
  block_sum<1>(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

struct FftState {
  int NX = 0, NY = 0;
  FFTPlan Px, Py;
  DevMem twx, twy, twxf, twyf, T, X, yf, mask, rg, U;
};
static FftState g_fft;

// mfista_amd_finalize has freed every allocation: forget the FFT-mode state
void reset_fft_state() { g_fft = FftState(); }

struct FftOp : Operator<double> {
  Context& C;
  FftState& F;
  explicit FftOp(Context& c, FftState& f) : C(c), F(f) {}
  int nky() const { return F.NY / 2 + 1; }
  int npart() const override { return elem_blocks(F.NX * nky()); }
  void fwd(const double* x, bool keep, double* part, const int* skip) override {
    launch_rowfft_fwd<double>(x, nullptr, nullptr, (double2*)F.T.p, F.NX, F.NY, F.NY, 0, F.Py,
                              (const double2*)F.twy.p, C.stream, skip);
    launch_colfft_fwd<double>((const double2*)F.T.p, (double2*)F.X.p, F.NX, nky(), F.NX, 0, 1, 1,
                              F.Px, (const double2*)F.twx.p, C.stream, skip);
    hipLaunchKernelGGL(k_fft_resid, dim3(npart()), dim3(MF_BLOCK), 0, C.stream,
                       (const double2*)F.X.p, (const double2*)F.yf.p, (const double*)F.mask.p,
                       keep ? (double2*)F.rg.p : nullptr, F.NX, nky(), F.NY,
                       1.0 / sqrt((double)F.NX * F.NY), part, skip);
  }
  void adj(double* g, const int* skip) override {
    launch_colfft_inv<double>((double2*)F.rg.p, (double2*)F.U.p, F.NX, nky(), F.NX, 0, 1, F.Px,
                              (const double2*)F.twx.p, C.stream, skip);
    launch_rowfft_c2r<double>((const double2*)F.U.p, nullptr, nullptr, 1.0, g, F.NX, F.NY, F.NY, 0,
                              F.Py, (const double2*)F.twy.p, C.stream, skip);
  }
};

int upload_twiddles_pub(const FFTPlan& P, DevMem& tf, DevMem& td, hipStream_t st);

// idx2mat + full2half (c/mfista_fft_lib.c:27-114) on the host, later entries
// overwriting earlier ones, then uploaded in the [ky][kx] layout
static int fft_setup(Context& C, const int* u_idx, const int* v_idx, const double* y_r,
                     const double* y_i, const double* noise, int M, int NX, int NY) {
  FftState& F = g_fft;
  if (NX < 2 || NY < 2 || (NX % 2) || (NY % 2) || NX > 2048 || NY > 2048) return MFISTA_E_SIZE;
  F.NX = NX;
  F.NY = NY;
  if (!make_fft_plan(NX, &F.Px) || !make_fft_plan(NY, &F.Py)) return MFISTA_E_SIZE;
  if (upload_twiddles_pub(F.Px, F.twxf, F.twx, C.stream) ||
      upload_twiddles_pub(F.Py, F.twyf, F.twy, C.stream))
    return MFISTA_E_HIP;
  const int nky = NY / 2 + 1;
  std::vector<double> mask((size_t)NX * nky, 0.0);
  std::vector<double2> yf((size_t)NX * nky, make_double2(0, 0));
  for (int k = 0; k < M; k++) {
    const int u = u_idx[k], v = v_idx[k];
    if (u < 0 || u >= NX || v < 0 || v >= NY) return MFISTA_E_ARG;
    if (v < nky) {
      const size_t id = (size_t)v * NX + u;
      mask[id] = 1 / noise[k];
      yf[id] = make_double2(y_r[k] / noise[k], y_i[k] / noise[k]);
    }
  }
  const size_t nh = (size_t)NX * nky;
  if (F.T.ensure(nh * 16) || F.X.ensure(nh * 16) || F.yf.ensure(nh * 16) || F.mask.ensure(nh * 8) ||
      F.rg.ensure(nh * 16) || F.U.ensure(nh * 16))
    return MFISTA_E_HIP;
  MF_CHECK(hipMemcpyAsync(F.mask.p, mask.data(), nh * 8, hipMemcpyHostToDevice, C.stream));
  MF_CHECK(hipMemcpyAsync(F.yf.p, yf.data(), nh * 16, hipMemcpyHostToDevice, C.stream));
  MF_CHECK(hipStreamSynchronize(C.stream));
  return 0;
}

static void fill_result(RESULT* R, int NX, int NY, double l1, double ltv, double ltsv, double chi2,
                        const double st[4]) {
  R->N = NX * NY;
  R->NX = NX;
  R->NY = NY;
  R->lambda_l1 = l1;
  R->lambda_tv = ltv;
  R->lambda_tsv = ltsv;
  R->sq_error = chi2;
  R->l1cost = st[0];
  R->N_active = (int)llround(st[3]);
  R->finalcost = chi2 / 2;
  if (l1 > 0) R->finalcost += l1 * st[0];
  if (ltsv > 0) {
    R->tsvcost = st[1];
    R->finalcost += ltsv * st[1];
  } else if (ltv > 0) {
    R->tvcost = st[2];
    R->finalcost += ltv * st[2];
  }
}

template <typename R>
static int op_value2(Context& C, Operator<R>& op, const R* x, double* twoF) {
  const int nb = op.npart();
  if (C.chi64.ensure((size_t)nb * 8)) return MFISTA_E_HIP;
  op.fwd(x, false, (double*)C.chi64.p, nullptr);
  std::vector<double> hpart(nb);
  MF_CHECK(hipMemcpyAsync(hpart.data(), C.chi64.p, nb * 8, hipMemcpyDeviceToHost, C.stream));
  MF_CHECK(hipStreamSynchronize(C.stream));
  double s = 0;
  for (int i = 0; i < nb; i++) s += hpart[i];
  *twoF = s;
  return 0;
}

static int mode_check(double ltv, double ltsv, int* tv_mode) {
  if (ltv == 0)
    *tv_mode = 0;
  else if (ltsv == 0)
    *tv_mode = 1;
  else {
    printf("You cannot set both of lambda_TV and lambda_TSV positive.\n");
    return MFISTA_E_BOTH_REG;
  }
  return 0;
}

int solve_fft_mode(Context& C, const int* u_idx, const int* v_idx, const double* y_r,
                   const double* y_i, const double* noise, int M, int NX, int NY, int maxiter,
                   double eps, double l1, double ltv, double ltsv, double cinit,
                   const double* xinit, double* xout, int nonneg, int box_flag,
                   const float* cl_box, RESULT* res, mfista_opts* opts) {
  int tv_mode, rc;
  if ((rc = mode_check(ltv, ltsv, &tv_mode))) return rc;
  if (M <= 0 || maxiter <= 0) return MFISTA_E_ARG;
  double eps_abs = 0;  // eps * mean |y|^2, c/mfista_fft_lib.c:715-717
  for (int i = 0; i < M; i++) eps_abs += y_r[i] * y_r[i] + y_i[i] * y_i[i];
  eps_abs *= eps / (double)M;
  if ((rc = fft_setup(C, u_idx, v_idx, y_r, y_i, noise, M, NX, NY))) return rc;
  const size_t nb = (size_t)NX * NY * 8;
  if ((rc = upload(C, C.h_x0, xinit, nb)) || (rc = upload(C, C.h_x1, nullptr, nb))) return rc;
  if (box_flag && (rc = upload(C, C.h_box, cl_box, (size_t)NX * NY * 4))) return rc;
  FftOp op(C, g_fft);
  auto t0 = std::chrono::steady_clock::now();
  RunStats rs;
  if (!tv_mode && nonneg != 0 && nonneg != 1) {
    printf("nonneg_flag must be chosen properly.\n");
    MF_CHECK(hipMemcpyAsync(C.h_x1.p, xout, nb, hipMemcpyHostToDevice, C.stream));
    rs.c = cinit;
  } else {
    RunParams P;
    P.NX = NX;
    P.NY = NY;
    P.maxiter = maxiter;
    P.nonneg = nonneg;
    P.box_flag = box_flag;
    P.tv_mode = tv_mode;
    P.fgp_iter = opts && opts->fgp_iter > 0 ? opts->fgp_iter : FGPITER;
    P.eps_abs = eps_abs;
    P.rel_tol = opts ? opts->rel_tol : 0.0;
    P.lambda_l1 = l1;
    P.lambda_tv = ltv;
    P.lambda_tsv = ltsv;
    P.cinit = cinit;
    P.eta_first = kEta;
    P.xinit = (const double*)C.h_x0.p;
    P.xout = (double*)C.h_x1.p;
    P.box = box_flag ? (const float*)C.h_box.p : nullptr;
    if ((rc = mfista_run<double>(C, op, P, &rs))) return rc;
  }
  auto t1 = std::chrono::steady_clock::now();
  double twoF = 0, st[4];
  if ((rc = op_value2<double>(C, op, (const double*)C.h_x1.p, &twoF))) return rc;
  if ((rc = image_stats64(C, (const double*)C.h_x1.p, NX, NY, st))) return rc;
  MF_CHECK(hipMemcpy(xout, C.h_x1.p, nb, hipMemcpyDeviceToHost));
  res->comp_time = std::chrono::duration<double>(t1 - t0).count();
  res->ITER = rs.ITER;
  res->nonneg = nonneg;
  res->Lip_const = rs.c;
  res->maxiter = maxiter;
  fill_result(res, NX, NY, l1, ltv, ltsv, twoF, st);
  res->M = M / 2;
  res->mean_sq_error = res->sq_error / (double)M;
  if (opts) {
    opts->max_trials = rs.max_trials;
    opts->total_trials = rs.total_trials;
    opts->solve_ms = res->comp_time * 1e3;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// DFT mode (c/mfista_dft_lib.c): A real M x N, column-major, FP64.
// r = y - A x (k_gemv_n, threads over rows, coalesced along each column);
// g = A^T r (k_gemv_t, one workgroup per column).

__global__ void __launch_bounds__(MF_BLOCK)
k_gemv_n(const double* __restrict__ A, const double* __restrict__ x, const double* __restrict__ y,
         double* __restrict__ r, int M, int N, double* part, const int* skip) {
  if (skip && *skip) return;
  __shared__ double xs[MF_BLOCK];
  __shared__ double red[16];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0;
  for (int j0 = 0; j0 < N; j0 += MF_BLOCK) {
    const int jj = j0 + threadIdx.x;
    xs[threadIdx.x] = (jj < N) ? x[jj] : 0.0;
    __syncthreads();
    const int jn = min(MF_BLOCK, N - j0);
    if (i < M)
      for (int j = 0; j < jn; j++) acc += A[(size_t)(j0 + j) * M + i] * xs[j];
    __syncthreads();
  }
  double v[1] = {0.0};
  if (i < M) {
    const double ri = y[i] - acc;
    r[i] = ri;
    v[0] = ri * ri;
  }
  block_sum<1>(v, red);
  if (threadIdx.x == 0) part[blockIdx.x] = v[0];
}

__global__ void __launch_bounds__(MF_BLOCK)
k_gemv_t(const double* __restrict__ A, const double* __restrict__ r, double* __restrict__ g, int M,
         int N, const int* skip) {
  if (skip && *skip) return;
  __shared__ double red[16];
  const int j = blockIdx.x;
  const double* a = A + (size_t)j * M;
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < M; i += blockDim.x) v[0] += a[i] * r[i];
  block_sum<1>(v, red);
  if (threadIdx.x == 0) g[j] = v[0];
}

struct DftOp : Operator<double> {
  Context& C;
  const double *A, *y;
  double *rk, *rt;
  int M, N;
  DftOp(Context& c, const double* A_, const double* y_, double* rk_, double* rt_, int M_, int N_)
      : C(c), A(A_), y(y_), rk(rk_), rt(rt_), M(M_), N(N_) {}
  int npart() const override { return elem_blocks(M); }
  void fwd(const double* x, bool keep, double* part, const int* skip) override {
    hipLaunchKernelGGL(k_gemv_n, dim3(npart()), dim3(MF_BLOCK), 0, C.stream, A, x, y, keep ? rk : rt,
                       M, N, part, skip);
  }
  void adj(double* g, const int* skip) override {
    hipLaunchKernelGGL(k_gemv_t, dim3(N), dim3(MF_BLOCK), 0, C.stream, A, rk, g, M, N, skip);
  }
};

int solve_dft_mode(Context& C, const double* y, const double* A, int M, int N, int NX, int NY,
                   int maxiter, double eps, double l1, double ltv, double ltsv, double cinit,
                   const double* xinit, double* xout, int nonneg, int box_flag,
                   const float* cl_box, RESULT* res, mfista_opts* opts) {
  memcpy(xout, xinit, sizeof(double) * (size_t)N);  // c/mfista_dft_lib.c:552
  int tv_mode, rc;
  if ((rc = mode_check(ltv, ltsv, &tv_mode))) return rc;
  if (M <= 0 || N <= 0 || N != NX * NY || maxiter <= 0) return MFISTA_E_ARG;
  const size_t nb = (size_t)N * 8;
  if ((rc = upload(C, C.h_A, A, (size_t)M * N * 8)) || (rc = upload(C, C.h_y, y, (size_t)M * 8)) ||
      (rc = upload(C, C.h_x0, xinit, nb)) || (rc = upload(C, C.h_x1, nullptr, nb)) ||
      (rc = upload(C, C.h_vr, nullptr, (size_t)M * 8)) || (rc = upload(C, C.h_vi, nullptr, (size_t)M * 8)))
    return rc;
  if (box_flag && (rc = upload(C, C.h_box, cl_box, (size_t)N * 4))) return rc;
  DftOp op(C, (const double*)C.h_A.p, (const double*)C.h_y.p, (double*)C.h_vr.p,
           (double*)C.h_vi.p, M, N);
  auto t0 = std::chrono::steady_clock::now();
  RunStats rs;
  rs.c = cinit;
  if (!tv_mode && nonneg != 0 && nonneg != 1) {
    printf("nonneg_flag must be chosen properly.\n");
  } else {
    RunParams P;
    P.NX = NX;
    P.NY = NY;
    P.maxiter = maxiter;
    P.nonneg = nonneg;
    P.box_flag = box_flag;
    P.tv_mode = tv_mode;
    P.fgp_iter = opts && opts->fgp_iter > 0 ? opts->fgp_iter : FGPITER;
    P.eps_abs = eps;  // unscaled, c/mfista_dft_lib.c:559
    P.rel_tol = opts ? opts->rel_tol : 0.0;
    P.lambda_l1 = l1;
    P.lambda_tv = ltv;
    P.lambda_tsv = ltsv;
    P.cinit = cinit;
    P.eta_first = kEta;
    P.xinit = (const double*)C.h_x0.p;
    P.xout = (double*)C.h_x1.p;
    P.box = box_flag ? (const float*)C.h_box.p : nullptr;
    if ((rc = mfista_run<double>(C, op, P, &rs))) return rc;
    MF_CHECK(hipMemcpy(xout, C.h_x1.p, nb, hipMemcpyDeviceToHost));
  }
  auto t1 = std::chrono::steady_clock::now();
  double twoF = 0, st[4];
  if ((rc = upload(C, C.h_x1, xout, nb))) return rc;
  op.rt = op.rk;  // residual of the final image kept for RESULT.residual
  if ((rc = op_value2<double>(C, op, (const double*)C.h_x1.p, &twoF))) return rc;
  if ((rc = image_stats64(C, (const double*)C.h_x1.p, NX, NY, st))) return rc;
  res->comp_time = std::chrono::duration<double>(t1 - t0).count();
  res->ITER = rs.ITER;
  res->nonneg = nonneg;
  res->Lip_const = rs.c;
  res->maxiter = maxiter;
  fill_result(res, NX, NY, l1, ltv, ltsv, twoF, st);
  res->M = M;
  res->N = N;
  res->mean_sq_error = res->sq_error / (double)M;
  res->residual = (double*)malloc(sizeof(double) * (size_t)M);
  if (!res->residual) return MFISTA_E_ARG;
  MF_CHECK(hipMemcpy(res->residual, op.rk, (size_t)M * 8, hipMemcpyDeviceToHost));
  if (opts) {
    opts->max_trials = rs.max_trials;
    opts->total_trials = rs.total_trials;
    opts->solve_ms = res->comp_time * 1e3;
  }
  return 0;
}

template void launch_tv_pre<float>(const float*, const float*, float*, int, const DevState*, hipStream_t, const int*);
template void launch_tv_pre<double>(const double*, const double*, double*, int, const DevState*, hipStream_t, const int*);

}  // namespace mf

using namespace mf;

extern "C" {

int mfista_imaging_core_fft_ex(const int* u_idx, const int* v_idx, const double* y_r,
                               const double* y_i, const double* noise_stdev, int M, int NX, int NY,
                               int maxiter, double eps, double lambda_l1, double lambda_tv,
                               double lambda_tsv, double cinit, const double* xinit, double* xout,
                               int nonneg_flag, int box_flag, const float* cl_box,
                               struct RESULT* res, mfista_opts* opts) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (!u_idx || !v_idx || !y_r || !y_i || !noise_stdev || !xinit || !xout || !res) return MFISTA_E_ARG;
  if (box_flag && !cl_box) return MFISTA_E_ARG;
  RESULT tmp = *res;
  int rc = solve_fft_mode(*C, u_idx, v_idx, y_r, y_i, noise_stdev, M, NX, NY, maxiter, eps, lambda_l1,
                          lambda_tv, lambda_tsv, cinit, xinit, xout, nonneg_flag, box_flag, cl_box,
                          &tmp, opts);
  if (rc) return rc;
  *res = tmp;
  return 0;
}

void mfista_imaging_core_fft(int* u_idx, int* v_idx, double* y_r, double* y_i,
                             double* noise_stdev, int M, int NX, int NY, int maxiter, double eps,
                             double lambda_l1, double lambda_tv, double lambda_tsv, double cinit,
                             double* xinit, double* xout, int nonneg_flag,
                             unsigned int fftw_plan_flag, int box_flag, float* cl_box,
                             struct RESULT* mfista_result) {
  (void)fftw_plan_flag;
  mfista_opts o;
  mfista_default_opts(&o);
  int rc = mfista_imaging_core_fft_ex(u_idx, v_idx, y_r, y_i, noise_stdev, M, NX, NY, maxiter, eps,
                                      lambda_l1, lambda_tv, lambda_tsv, cinit, xinit, xout,
                                      nonneg_flag, box_flag, cl_box, mfista_result, &o);
  if (rc && rc != MFISTA_E_BOTH_REG)
    fprintf(stderr, "[mfista_amd] mfista_imaging_core_fft failed with status %d\n", rc);
}

int mfista_imaging_core_dft_ex(const double* y, const double* A, int M, int N, int NX, int NY,
                               int maxiter, double eps, double lambda_l1, double lambda_tv,
                               double lambda_tsv, double cinit, const double* xinit, double* xout,
                               int nonneg_flag, int box_flag, const float* cl_box,
                               struct RESULT* res, mfista_opts* opts) {
  std::lock_guard<std::mutex> lock(global_mutex());
  Context* C = get_context();
  if (!C) return MFISTA_E_HIP;
  if (!y || !A || !xinit || !xout || !res) return MFISTA_E_ARG;
  if (box_flag && !cl_box) return MFISTA_E_ARG;
  RESULT tmp = *res;
  int rc = solve_dft_mode(*C, y, A, M, N, NX, NY, maxiter, eps, lambda_l1, lambda_tv, lambda_tsv,
                          cinit, xinit, xout, nonneg_flag, box_flag, cl_box, &tmp, opts);
  if (rc) return rc;
  *res = tmp;
  return 0;
}

void mfista_imaging_core_dft(double* y, double* A, int* M, int* N, int NX, int NY, int maxiter,
                             double eps, double lambda_l1, double lambda_tv, double lambda_tsv,
                             double cinit, double* xinit, double* xout, int nonneg_flag,
                             int looe_flag, int box_flag, float* cl_box,
                             struct RESULT* mfista_result) {
  (void)looe_flag;
  mfista_opts o;
  mfista_default_opts(&o);
  int rc = mfista_imaging_core_dft_ex(y, A, *M, *N, NX, NY, maxiter, eps, lambda_l1, lambda_tv,
                                      lambda_tsv, cinit, xinit, xout, nonneg_flag, box_flag, cl_box,
                                      mfista_result, &o);
  if (rc && rc != MFISTA_E_BOTH_REG)
    fprintf(stderr, "[mfista_amd] mfista_imaging_core_dft failed with status %d\n", rc);
}

}  // extern "C"
