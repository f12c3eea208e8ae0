// persistent.hip -- the whole NUFFT MFISTA solve in one cooperative launch.
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
#include <cmath>
#include <cstring>
#include <vector>

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
  double2 *T64, *Gf64;
  double2* vm[3];  // weighted model vectors W A (.), roles rotate on acceptance
  const int4* chunks;
  int nchunks;
  const unsigned* ent;
  double *x, *xnew, *z;  // iterates in FP64 (FP32 iterates stagnate near a tight tolerance)
  float* gr;
  const float* box;
  float *rA, *sA, *rB, *sB, *pp, *qq;
  const double* fgp_c;  // c1[k] = fgp_c[2k], c2[k] = fgp_c[2k+1]
  int fgp_iter, tv_mode, maxiter, nonneg;
  double lambda_l1, lambda_tsv, lambda_tv, eps_abs, rel_tol, eta_first, cinit, adj_scale;
  double* part;
  int n_deg, n_row, n_ky, n_elem;
  double* cost;
  DevState* S;
  unsigned long long* bar;  // [0] arrival counter, [1] error flag
};

// Grid-wide barrier for a cooperative launch.  bar[0] is a monotonic 64-bit
// arrival counter: the k-th barrier is passed once it reaches k * nblk, so no
// reset (and no ordering of a reset) is needed.  Each workgroup pays exactly
// one agent-scope release fence (L2 write-back) before arriving and one
// acquire fence (L2 invalidate) after the wait; the wait itself polls with
// relaxed atomic loads, which leave the caches alone.  A bounded spin sets
// the error flag bar[1] instead of waiting for ever.
__device__ inline bool grid_sync(unsigned long long* bar, unsigned nblk, unsigned long long& epoch) {
  __syncthreads();
  __shared__ int fail;
  epoch += 1;
  if (threadIdx.x == 0) {
    fail = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(&bar[0], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned long long target = epoch * (unsigned long long)nblk;
    long spins = 0;
    while (__hip_atomic_load(&bar[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      __builtin_amdgcn_s_sleep(1);
      if ((++spins & 1023) == 0 &&
          (__hip_atomic_load(&bar[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
           spins > (1L << 26))) {
        __hip_atomic_store(&bar[1], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        fail = 1;
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return fail == 0;
}

__device__ inline double blk_reduce(const double* p, int cnt, int stride, double* red) {
  double v[1] = {0.0};
  for (int i = threadIdx.x; i < cnt; i += blockDim.x) v[0] += p[(size_t)i * stride];
  block_sum<1>(v, red);
  return v[0];
}

struct BlockState {
  double c, mu, costtmp, l1_x, mom_a, mom_b;
  int iter, trial, accept, mom_pending, stop, ITER, max_trials, total_trials, status, prologue;
  int ix, in, iz;  // roles of the model vectors: W A x, W A xnew, W A z
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

constexpr int kT1W = 2;  // per-row-task trial partials: |d|^2, |xnew|_1

template <int G>
__global__ void __launch_bounds__(MF_BLOCK) k_mfista_persistent(PArgs A) {
  extern __shared__ double4 smem_d4[];
  __shared__ double red[16 * kProxW];
  __shared__ BlockState bs;
  __shared__ double ring[kTD + 1];
  float2* sm32 = reinterpret_cast<float2*>(smem_d4);
  double2* sm64 = reinterpret_cast<double2*>(smem_d4);
  const GridGeom& g = A.g;
  const unsigned nb = gridDim.x;
  const int NX = g.NX, NY = g.NY, n = NX * NY, W = NX, H = NY;
  const bool tsv = A.lambda_tsv > 0.0;
  const bool tv = A.tv_mode != 0;
  unsigned long long epoch = 0;  // barriers passed by this workgroup
  // partial-sum regions
  double* P_chix = A.part;
  double* P_lip = P_chix + A.n_deg;
  double* P_t1 = P_lip + A.n_deg;                 // kT1W per row task
  double* P_reg = P_t1 + kT1W * A.n_row;          // 2 per element task
  double* P_init = P_reg + 2 * A.n_elem;          // 4 per row task

  // ---------------- initial cost and models: W A x0 -> vm[0] (x), vm[2] (z)
  for (int t = blockIdx.x; t < A.n_row; t += nb) {
    double st[4] = {0, 0, 0, 0};
    auto load = [&](int i, int j) -> double {
      const int f = i * NY + j;
      const double xc = A.x[f];
      A.z[f] = xc;
      st[0] += fabs(xc);
      st[3] += (xc != 0.0) ? 1.0 : 0.0;
      const int c = f % W, r = f / W;
      double dh = 0, dv = 0;
      if (c < W - 1) dh = xc - A.x[f + 1];
      if (r < H - 1) dv = xc - A.x[f + W];
      st[1] += dh * dh + dv * dv;
      if (c < W - 1 && r < H - 1)
        st[2] += sqrt(dh * dh + dv * dv);
      else if (c < W - 1)
        st[2] += fabs(dh);
      else if (r < H - 1)
        st[2] += fabs(dv);
      return xc * A.e4xd[i] * A.e4yd[j];
    };
    st_rowfft_fwd<double>(t, 1, NX, NY, g.Ly, NY / 2, A.Py, A.twyd, A.T64, sm64, load);
    block_sum<4>(st, red);
    if (threadIdx.x == 0)
      for (int q = 0; q < 4; q++) P_init[4 * t + q] = st[q];
  }
  if (!grid_sync(A.bar, nb, epoch)) return;
  for (int t = blockIdx.x; t < A.n_ky; t += nb)
    st_colfft_fwd<double>(t, 1, A.T64, A.Gf64, NX, NY + 1, g.Lx, NX / 2, 0, g.MSP, A.Px, A.twxd, sm64);
  if (!grid_sync(A.bar, nb, epoch)) return;
  for (int t = blockIdx.x; t < A.n_deg; t += nb) {
    double v[2];
    st_degrid_model<G>(t, g, A.pd, A.Gf64, nullptr, A.vm[0], A.vm[2], v);
    block_sum<2>(v, red);
    if (threadIdx.x == 0) P_chix[t] = v[0];
  }
  if (!grid_sync(A.bar, nb, epoch)) return;
  {
    const double chi = blk_reduce(P_chix, A.n_deg, 1, red);
    const double l1 = blk_reduce(P_init + 0, A.n_row, 4, red);
    const double ts = blk_reduce(P_init + 1, A.n_row, 4, red);
    const double tvs = blk_reduce(P_init + 2, A.n_row, 4, red);
    if (threadIdx.x == 0) {
      double c0 = chi / 2 + A.lambda_l1 * l1;
      if (tv)
        c0 += A.lambda_tv * tvs;
      else if (tsv)
        c0 += A.lambda_tsv * ts;
      bs.c = A.cinit;
      bs.mu = 1.0;
      bs.costtmp = c0;
      bs.l1_x = l1;
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
      const double2* mx = A.vm[bs.ix];
      const double2* mn = A.vm[bs.in];
      double2* mz = A.vm[bs.iz];
      if (mom) {
        for (int t = blockIdx.x; t < A.n_elem; t += nb) {
          const int f = t * blockDim.x + threadIdx.x;
          if (f < n) {
            const double xf = A.x[f], xn = A.xnew[f];
            A.z[f] = ma * xf + mb * xn;
            if (acc) A.x[f] = xn;
          }
        }
        for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < g.M; s += nb * blockDim.x) {
          const double2 a = mx[s], b = mn[s];
          mz[s] = make_double2(ma * a.x + mb * b.x, ma * a.y + mb * b.y);
        }
      }
      auto resid = [&](int s) -> float2 {
        const double w = A.pd.w[s];
        const double2 V = A.pd.vis[s];
        double2 m;
        if (mom) {
          const double2 a = mx[s], b = mn[s];
          m = make_double2(ma * a.x + mb * b.x, ma * a.y + mb * b.y);
        } else {
          m = mz[s];
        }
        return make_float2((float)(V.x * w - m.x), (float)(V.y * w - m.y));
      };
      for (int t = blockIdx.x; t < A.nchunks; t += nb)
        st_grid<float>(t, g, A.pf, resid, A.chunks, A.ent, A.Hg, sm32);
      if (!grid_sync(A.bar, nb, epoch)) return;
      if (threadIdx.x == 0 && mom && acc) {  // W A x <- W A xnew: swap roles
        const int tmp = bs.ix;
        bs.ix = bs.in;
        bs.in = tmp;
      }
      // P2: inverse column pass (clears Hg behind it)
      for (int t = blockIdx.x; t < A.n_ky; t += nb)
        st_colfft_inv<float>(t, 1, A.Hg, A.U, NX, NY + 1, g.Lx, NX / 2, 0, A.Px, A.twxf, sm32);
      if (!grid_sync(A.bar, nb, epoch)) return;
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
      float gv = v * sc * A.e4xf[i] * A.e4yf[j];
      if (tsv) {
        const int cc = f % W, r = f / W;
        const double zc = A.z[f];
        double d = 0.0;
        if (cc < W - 1) d += zc - A.z[f + 1];
        if (cc > 0) d += zc - A.z[f - 1];
        if (r < H - 1) d += zc - A.z[f + W];
        if (r > 0) d += zc - A.z[f - W];
        gv = gv - lt * (float)(2.0 * d);
      }
      A.gr[f] = gv;
    };
    if (tv) {
      if (trial_first) {
        for (int t = blockIdx.x; t < A.n_row; t += nb)
          st_rowfft_c2r<float>(t, 1, A.U, NX, NY, g.Ly, NY / 2, A.Py, A.twyf, sm32, gstore);
        if (!grid_sync(A.bar, nb, epoch)) return;
      }
      // FGP iterations (c/mfista_TV_lib.c:189-384), b = z + g/c (nonneg: -l1/c + g/c + z)
      const float bshift = (A.nonneg == 1) ? (float)(-A.lambda_l1 / c) : 0.f;
      const float a3 = 1.f / (8.f * lam);
      float *ri = A.rA, *si = A.sA, *ro = A.rB, *so = A.sB;
      for (int k = 0; k < A.fgp_iter; k++) {
        const float c1 = (float)A.fgp_c[2 * k], c2 = (float)A.fgp_c[2 * k + 1];
        const int first = (k == 0);
        for (int t = blockIdx.x; t < A.n_elem; t += nb) {
          const int f = t * blockDim.x + threadIdx.x;
          if (f >= n) continue;
          const int i = f % W, j = f / W;
          auto xat = [&](int ii, int jj) -> float {
            const int ff = jj * W + ii;
            const float bb = bshift + cinv * A.gr[ff] + (float)A.z[ff];
            const float L = first ? 0.f : lpq_f(ri, si, ii, jj, W, H);
            return fgp_proj(bb - lam * L, eta, A.nonneg == 1, A.box && A.box[ff] == 0.0f);
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
            const float pold = first ? 0.f : A.pp[kk];
            ro[kk] = np_ * c1 + c2 * pold;
            A.pp[kk] = np_;
          }
          if (hq) {
            const int kk = W * j + i;
            const float qold = first ? 0.f : A.qq[kk];
            so[kk] = nq_ * c1 + c2 * qold;
            A.qq[kk] = nq_;
          }
        }
        if (!grid_sync(A.bar, nb, epoch)) return;
        float* tmp = ri;
        ri = ro;
        ro = tmp;
        tmp = si;
        si = so;
        so = tmp;
      }
    }
    // T1: [c2r + gradient on the first trial] -> xnew (prox or final FGP step),
    // d = xnew - z (FP64), partials, FP64 forward row pass of d
    for (int t = blockIdx.x; t < A.n_row; t += nb) {
      if (!tv && trial_first)
        st_rowfft_c2r<float>(t, 1, A.U, NX, NY, g.Ly, NY / 2, A.Py, A.twyf, sm32, gstore);
      double v[kT1W] = {0, 0};
      const double cinv_d = 1.0 / c, eta_d = A.lambda_l1 / c, lam_d = A.lambda_tv / c;
      const double bshift = (tv && A.nonneg == 1) ? -eta_d : 0.0;
      const bool have_dual = tv && A.fgp_iter > 0;
      auto load = [&](int i, int j) -> double {
        const int f = i * NY + j;
        const double zf = A.z[f];
        const bool boxed = A.box && A.box[f] == 0.0f;
        double xn;
        if (!tv) {
          xn = prox_p<double>((double)A.gr[f] * cinv_d + zf, eta_d, A.nonneg, boxed);
        } else {
          const int ii = f % W, jj = f / W;
          const double bb = bshift + cinv_d * (double)A.gr[f] + zf;
          const double L = have_dual ? (double)lpq_f(A.pp, A.qq, ii, jj, W, H) : 0.0;
          xn = fgp_proj_d(bb - lam_d * L, eta_d, A.nonneg == 1, boxed);
        }
        A.xnew[f] = xn;
        const double d = xn - zf;
        v[0] += d * d;
        v[1] += fabs(xn);
        return d * A.e4xd[i] * A.e4yd[j];
      };
      st_rowfft_fwd<double>(t, 1, NX, NY, g.Ly, NY / 2, A.Py, A.twyd, A.T64, sm64, load);
      block_sum<kT1W>(v, red);
      if (threadIdx.x == 0)
        for (int q = 0; q < kT1W; q++) P_t1[kT1W * t + q] = v[q];
    }
    if (!grid_sync(A.bar, nb, epoch)) return;
    // T2: FP64 column pass of d; regulariser partials of xnew and d
    for (int t = blockIdx.x; t < A.n_ky; t += nb)
      st_colfft_fwd<double>(t, 1, A.T64, A.Gf64, NX, NY + 1, g.Lx, NX / 2, 0, g.MSP, A.Px, A.twxd, sm64);
    for (int t = blockIdx.x; t < A.n_elem; t += nb) {
      double r[2] = {0, 0};
      const int f = t * blockDim.x + threadIdx.x;
      if (f < n) reg_terms(A.xnew, A.z, f, W, H, tv, r);
      block_sum<2>(r, red);
      if (threadIdx.x == 0) {
        P_reg[2 * t] = r[0];
        P_reg[2 * t + 1] = r[1];
      }
    }
    if (!grid_sync(A.bar, nb, epoch)) return;
    // T3: FP64 degrid of d: W A xnew = W A z + W A d, chi^2(xnew), |W A d|^2
    {
      const double2* mz = A.vm[bs.iz];
      double2* mn = A.vm[bs.in];
      for (int t = blockIdx.x; t < A.n_deg; t += nb) {
        double v[2];
        st_degrid_model<G>(t, g, A.pd, A.Gf64, mz, mn, nullptr, v);
        block_sum<2>(v, red);
        if (threadIdx.x == 0) {
          P_chix[t] = v[0];
          P_lip[t] = v[1];
        }
      }
    }
    if (!grid_sync(A.bar, nb, epoch)) return;

    // ---------------- control (every workgroup, identical arithmetic)
    const double chi_x = blk_reduce(P_chix, A.n_deg, 1, red);
    const double lip = blk_reduce(P_lip, A.n_deg, 1, red);
    const double s_d2 = blk_reduce(P_t1 + 0, A.n_row, kT1W, red);
    const double l1n = blk_reduce(P_t1 + 1, A.n_row, kT1W, red);
    const double regn = blk_reduce(P_reg + 0, A.n_elem, 2, red);
    const double tsvd = blk_reduce(P_reg + 1, A.n_elem, 2, red);
    if (threadIdx.x == 0) {
      const int it = bs.iter;
      if (prologue) {
        ring[it % (kTD + 1)] = bs.costtmp;
        if (blockIdx.x == 0) A.cost[it] = bs.costtmp;
        bs.prologue = 0;
      }
      double F = chi_x / 2 + ((!tv && tsv) ? A.lambda_tsv * regn : 0.0);
      const double lhs = lip / 2 + ((!tv && tsv) ? A.lambda_tsv * tsvd : 0.0);
      const double rhs = bs.c * s_d2 / 2;
      bs.trial += 1;
      bs.total_trials += 1;
      const double etaf = (it == 0) ? A.eta_first : kEta;
      const long cap = (long)A.maxiter * 64 + 100000;
      if (bs.total_trials > cap) {  // cannot happen for a convergent run; bounds the launch
        bs.status = 2;
        bs.stop = 1;
        bs.ITER = it + 1;
        bs.mom_pending = 0;
      } else if (!isfinite(F) || !isfinite(lhs)) {
        bs.status = 1;
        bs.stop = 1;
        bs.ITER = it + 1;
        bs.mom_pending = 0;
      } else {
        const bool ok = lhs <= rhs;
        if (!ok && bs.trial < A.maxiter) {
          bs.c *= etaf;
        } else {
          if (!ok) bs.c *= etaf;
          if (bs.trial > bs.max_trials) bs.max_trials = bs.trial;
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
            if (it > 1 && bs.l1_x == 0) stop = true;
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
        }
      }
    }
    __syncthreads();
    if (bs.stop) break;
  }

  // the accepted image of the last iteration (the reference's x = xnew)
  if (bs.mom_pending && bs.accept) {
    for (int t = blockIdx.x; t < A.n_elem; t += nb) {
      const int f = t * blockDim.x + threadIdx.x;
      if (f < n) A.x[f] = A.xnew[f];
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    DevState* S = A.S;
    S->c = bs.c;
    S->mu = bs.mu;
    S->costtmp = bs.costtmp;
    S->ITER = bs.ITER;
    S->iter = bs.iter;
    S->max_trials = bs.max_trials;
    S->total_trials = bs.total_trials;
    S->status = bs.status;
    S->done = 1;
  }
}

int solve_nufft_persistent(Context& C, const RunParams& P, RunStats* out) {
  NufftState& N = C.nu;
  hipStream_t st = C.stream;
  const GridGeom& g = N.geom;
  const int NX = P.NX, NY = P.NY, n = NX * NY;
  const int n_row = (NX + 1) / 2, n_ky = NY + 1, n_elem = elem_blocks(n);
  const int n_deg = degrid_blocks(N.M, N.MSP);
  const size_t npart = 2 * (size_t)n_deg + (kT1W + 4) * (size_t)n_row + 2 * (size_t)n_elem;
  const size_t rb = (size_t)n * 4;
  if (C.px.ensure(2 * rb) || C.pz.ensure(2 * rb) || C.pxn.ensure(2 * rb) || C.g.ensure(rb) ||
      C.vmod.ensure(3 * (size_t)N.M * sizeof(double2)) ||
      C.part.ensure(npart * 8) || C.state.ensure(sizeof(DevState)) ||
      C.cost.ensure((size_t)(P.maxiter > 0 ? P.maxiter : 1) * 8) || C.bar.ensure(64) ||
      C.fgpc.ensure((size_t)2 * (P.fgp_iter > 0 ? P.fgp_iter : 1) * 8))
    return MFISTA_E_HIP;
  if (P.tv_mode && C.fgp.ensure(7 * rb)) return MFISTA_E_HIP;
  if (nufft_ensure64(C)) return MFISTA_E_HIP;
  {
    std::vector<double> cc(2 * (size_t)(P.fgp_iter > 0 ? P.fgp_iter : 1), 0.0);
    double t = 1;
    for (int k = 0; k < P.fgp_iter; k++) {
      const double tn = (1 + sqrt(1 + 4 * t * t)) / 2;
      cc[2 * k] = 1 + (t - 1) / tn;
      cc[2 * k + 1] = -(t - 1) / tn;
      t = tn;
    }
    MF_CHECK(hipMemcpyAsync(C.fgpc.p, cc.data(), cc.size() * 8, hipMemcpyHostToDevice, st));
  }
  MF_CHECK(hipMemcpyAsync(C.px.p, P.xinit, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  MF_CHECK(hipMemsetAsync(C.bar.p, 0, 64, st));
  PArgs A;
  memset(&A, 0, sizeof(A));
  A.g = g;
  A.Px = N.Px;
  A.Py = N.Py;
  A.pf = N.pf();
  A.pd = N.pd();
  A.twxf = (const float2*)N.twxf.p;
  A.twyf = (const float2*)N.twyf.p;
  A.twxd = (const double2*)N.twxd.p;
  A.twyd = (const double2*)N.twyd.p;
  A.e4xf = (const float*)N.e4xf.p;
  A.e4yf = (const float*)N.e4yf.p;
  A.e4xd = (const double*)N.e4xd.p;
  A.e4yd = (const double*)N.e4yd.p;
  A.Hg = (float2*)N.Hg.p;
  A.U = (float2*)N.U.p;
  A.T64 = (double2*)N.T64.p;
  A.Gf64 = (double2*)N.Gf64.p;
  for (int q = 0; q < 3; q++) A.vm[q] = (double2*)C.vmod.p + (size_t)q * N.M;
  A.chunks = (const int4*)N.chunks.p;
  A.nchunks = N.nchunks;
  A.ent = (const unsigned*)N.ent.p;
  A.x = (double*)C.px.p;
  A.z = (double*)C.pz.p;
  A.xnew = (double*)C.pxn.p;
  A.gr = (float*)C.g.p;
  A.box = P.box_flag ? P.box : nullptr;
  if (P.tv_mode) {
    float* fb = (float*)C.fgp.p;
    A.rA = fb;
    A.sA = fb + n;
    A.rB = fb + 2 * (size_t)n;
    A.sB = fb + 3 * (size_t)n;
    A.pp = fb + 4 * (size_t)n;
    A.qq = fb + 5 * (size_t)n;
  }
  A.fgp_c = (const double*)C.fgpc.p;
  A.fgp_iter = P.tv_mode ? P.fgp_iter : 0;
  A.tv_mode = P.tv_mode;
  A.maxiter = P.maxiter;
  A.nonneg = P.nonneg;
  A.lambda_l1 = P.lambda_l1;
  A.lambda_tsv = P.tv_mode ? 0.0 : P.lambda_tsv;
  A.lambda_tv = P.tv_mode ? P.lambda_tv : 0.0;
  A.eps_abs = P.eps_abs;
  A.rel_tol = P.rel_tol;
  A.eta_first = P.eta_first;
  A.cinit = P.cinit;
  A.adj_scale = 1.0 / (8.0 * (double)NX * NY);
  A.part = (double*)C.part.p;
  A.n_deg = n_deg;
  A.n_row = n_row;
  A.n_ky = n_ky;
  A.n_elem = n_elem;
  A.cost = (double*)C.cost.p;
  A.S = (DevState*)C.state.p;
  A.bar = (unsigned long long*)C.bar.p;
  // dynamic LDS: the largest stage (FP64 row/column FFT or the gridding tile copies)
  const int L = g.Lx > g.Ly ? g.Lx : g.Ly;
  size_t sh = 2 * (size_t)L * sizeof(double2);
  const size_t shg = 4 * (size_t)g.TX * g.TY * sizeof(float2);
  if (shg > sh) sh = shg;
  const void* fn = (N.MSP * 2 <= 16) ? (const void*)k_mfista_persistent<16>
                                      : (const void*)k_mfista_persistent<32>;
  int per_cu = 0, ncu = 0, dev = 0;
  MF_CHECK(hipGetDevice(&dev));
  MF_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  MF_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, MF_BLOCK, sh));
  if (per_cu < 1) return MFISTA_E_HIP;
  const int want = C.persist_blocks_per_cu > 0 ? C.persist_blocks_per_cu : 1;
  const int nblk = ncu * (per_cu < want ? per_cu : want);
  void* args[] = {&A};
  hipEvent_t e0, e1;
  MF_CHECK(hipEventCreate(&e0));
  MF_CHECK(hipEventCreate(&e1));
  MF_CHECK(hipEventRecord(e0, st));
  MF_CHECK(hipLaunchCooperativeKernel(fn, dim3(nblk), dim3(MF_BLOCK), args, (unsigned)sh, st));
  MF_CHECK(hipEventRecord(e1, st));
  MF_CHECK(hipStreamSynchronize(st));
  float kms = 0;
  MF_CHECK(hipEventElapsedTime(&kms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  out->kernel_ms = kms;
  C.persist_blocks = nblk;
  MF_CHECK(hipGetLastError());
  unsigned long long hb[2];
  MF_CHECK(hipMemcpy(hb, C.bar.p, sizeof(hb), hipMemcpyDeviceToHost));
  if (hb[1]) {
    fprintf(stderr, "[mfista_amd] persistent solver: grid barrier timed out\n");
    return MFISTA_E_HIP;
  }
  DevState h;
  MF_CHECK(hipMemcpy(&h, C.state.p, sizeof(h), hipMemcpyDeviceToHost));
  out->ITER = h.ITER;
  out->c = h.c;
  out->max_trials = h.max_trials;
  out->total_trials = h.total_trials;
  out->status = h.status;
  out->steps = h.total_trials;
  MF_CHECK(hipMemcpyAsync(P.xout, C.px.p, (size_t)n * 8, hipMemcpyDeviceToDevice, st));
  MF_CHECK(hipStreamSynchronize(st));
  return 0;
}

}  // namespace mf
