/*
 * mfista_oracle.c -- FP64 CPU restatement of the reference MFISTA imaging library.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the timed CPU baseline.  The
 * product (sparseimaging_amd) never links or calls this file.
 *
 * Reference (ikeda46/sparseimaging, GPL-3) is restated, not copied:
 *   c/mfista_nufft_lib.c, c/mfista_fft_lib.c, c/mfista_dft_lib.c,
 *   c/mfista_tools.c, c/mfista_TV_lib.c, c/mfista.h,
 *   c++/mfista_nufft_lib.cpp, c++/mfista_tools.cpp, c++/mfista.hpp.
 * FFTW and BLAS are replaced by the textbook operations they compute: an
 * unnormalised mixed-radix DFT with FFTW's sign and half-plane conventions,
 * and plain loops for dcopy/daxpy/ddot/dasum/dgemv.
 *
 * Built twice: serially (liboracle_mfista.so, the checker) and with OpenMP
 * (liboracle_mfista_omp.so, the all-cores CPU baseline, as the C++ reference
 * is built with -fopenmp, c++/makefile:23-24): the 2-D FFT passes, the
 * degridding and the 12-tap gridding (per-thread grids summed afterwards) then
 * run in parallel.  The #pragma lines are inert in the serial build.
 *
 * Deliberate deviations (documented): a tap that would leave the 2N x 2N grid
 * is rejected up front (the C library writes out of bounds, the C++ one warns);
 * the C++ soft_threshold_box bug (threshold ignored when box_flag=1) is not
 * reproduced; the caller's arrays are never permuted.
 */
#include "mfista_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

void or_free(void *p) { free(p); }

void or_default_opts(or_opts *o) {
  memset(o, 0, sizeof(*o));
  o->preset = OR_PRESET_C24;
  o->fgp_iter = OR_FGPITER;
}

static double *dvec(size_t n) {
  double *p = (double *)calloc(n ? n : 1, sizeof(double));
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

/* ======================================================================
 * Mixed-radix DFT.  Computes X[k] = sum_j x[j] exp(sign*2*pi*i*j*k/n),
 * unnormalised, the definition FFTW's plans execute.
 * ====================================================================== */

typedef struct {
  int n, nf, f[64];
  double *tw; /* tw[2j], tw[2j+1] = cos, sin of -2 pi j / n */
} or_fftplan;

#define OR_NPLANS 24
static or_fftplan g_plans[OR_NPLANS];
static int g_nplans = 0;

static const or_fftplan *fft_plan(int n) {
  for (int i = 0; i < g_nplans; i++)
    if (g_plans[i].n == n) return &g_plans[i];
  or_fftplan *P;
  if (g_nplans < OR_NPLANS) {
    P = &g_plans[g_nplans++];
  } else { /* recycle the oldest slot */
    P = &g_plans[0];
    free(P->tw);
  }
  P->n = n;
  P->nf = 0;
  int m = n;
  while (m % 4 == 0) { P->f[P->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { P->f[P->nf++] = 2; m /= 2; }
  for (int p = 3; m > 1; p += 2)
    while (m % p == 0) { P->f[P->nf++] = p; m /= p; }
  P->tw = dvec(2 * (size_t)n);
  for (int j = 0; j < n; j++) {
    double a = -2.0 * M_PI * (double)j / (double)n;
    P->tw[2 * j] = cos(a);
    P->tw[2 * j + 1] = sin(a);
  }
  return P;
}

/* out[0..n) = DFT_n of in[0], in[is], in[2 is], ...; twiddle step into the
 * root table of the full-length plan is twstep = N0/n. */
static void fft_rec(const or_fftplan *P, int n, const double *in, int is, double *out,
                    const int *f, int sign, int twstep) {
  if (n == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  const int p = f[0], m = n / p, N0 = P->n;
  for (int r = 0; r < p; r++)
    fft_rec(P, m, in + 2 * (size_t)r * is, is * p, out + 2 * (size_t)r * m, f + 1, sign,
            twstep * p);
  double tr[64], ti[64];
  double *gr = tr, *gi = ti;
  if (p > 64) {
    gr = dvec(p);
    gi = dvec(p);
  }
  const double sg = (sign > 0) ? -1.0 : 1.0; /* conjugate the forward table for sign>0 */
  for (int k = 0; k < m; k++) {
    for (int r = 0; r < p; r++) {
      double xr = out[2 * ((size_t)r * m + k)], xi = out[2 * ((size_t)r * m + k) + 1];
      if (r == 0 || k == 0) {
        gr[r] = xr;
        gi[r] = xi;
      } else {
        long e = ((long)r * k * twstep) % N0;
        double wr = P->tw[2 * e], wi = sg * P->tw[2 * e + 1];
        gr[r] = xr * wr - xi * wi;
        gi[r] = xr * wi + xi * wr;
      }
    }
    if (p == 2) {
      out[2 * k] = gr[0] + gr[1];
      out[2 * k + 1] = gi[0] + gi[1];
      out[2 * (m + k)] = gr[0] - gr[1];
      out[2 * (m + k) + 1] = gi[0] - gi[1];
    } else if (p == 4) {
      /* w = exp(sign*2*pi*i/4) = sign*i */
      double ar = gr[0] + gr[2], ai = gi[0] + gi[2];
      double br = gr[1] + gr[3], bi = gi[1] + gi[3];
      double cr = gr[0] - gr[2], ci = gi[0] - gi[2];
      double dr = gr[1] - gr[3], di = gi[1] - gi[3];
      /* w*d = sign*i*(dr + i di) = sign*(-di + i dr) */
      double wdr = -(double)sign * di, wdi = (double)sign * dr;
      out[2 * k] = ar + br;
      out[2 * k + 1] = ai + bi;
      out[2 * (2 * m + k)] = ar - br;
      out[2 * (2 * m + k) + 1] = ai - bi;
      out[2 * (m + k)] = cr + wdr;
      out[2 * (m + k) + 1] = ci + wdi;
      out[2 * (3 * m + k)] = cr - wdr;
      out[2 * (3 * m + k) + 1] = ci - wdi;
    } else {
      for (int q = 0; q < p; q++) {
        double sr = 0, si = 0;
        for (int r = 0; r < p; r++) {
          long e = ((long)r * q % p) * (long)m * twstep % N0;
          double wr = P->tw[2 * e], wi = sg * P->tw[2 * e + 1];
          sr += gr[r] * wr - gi[r] * wi;
          si += gr[r] * wi + gi[r] * wr;
        }
        out[2 * ((size_t)q * m + k)] = sr;
        out[2 * ((size_t)q * m + k) + 1] = si;
      }
    }
  }
  if (p > 64) {
    free(gr);
    free(gi);
  }
}

void or_fft_c2c(int n, double *data, int sign) {
  if (n <= 1) return;
  const or_fftplan *P = fft_plan(n);
  double *tmp = dvec(2 * (size_t)n);
  fft_rec(P, n, data, 1, tmp, P->f, sign, 1);
  memcpy(data, tmp, sizeof(double) * 2 * (size_t)n);
  free(tmp);
}

/* 2-D real-to-half-complex, FFTW r2c semantics: out[i*h+k], h = n1/2+1,
 * k in [0, n1/2], i in [0, n0).  Rows are transformed two at a time. */
void or_fft2_r2c(int n0, int n1, const double *in, double *out) {
  const int h = n1 / 2 + 1;
  (void)fft_plan(n0); /* plans are made here, before any parallel loop reads them */
  (void)fft_plan(n1);
#pragma omp parallel
  {
  double *z = dvec(2 * (size_t)n1);
#pragma omp for schedule(static)
  for (int i = 0; i < n0; i += 2) {
    const int two = (i + 1 < n0);
    for (int j = 0; j < n1; j++) {
      z[2 * j] = in[(size_t)i * n1 + j];
      z[2 * j + 1] = two ? in[(size_t)(i + 1) * n1 + j] : 0.0;
    }
    or_fft_c2c(n1, z, -1);
    for (int k = 0; k < h; k++) {
      int kk = (n1 - k) % n1;
      double zr = z[2 * k], zi = z[2 * k + 1], cr = z[2 * kk], ci = -z[2 * kk + 1];
      /* A = (Z[k] + conj Z[-k])/2, B = (Z[k] - conj Z[-k])/(2i) */
      out[2 * ((size_t)i * h + k)] = 0.5 * (zr + cr);
      out[2 * ((size_t)i * h + k) + 1] = 0.5 * (zi + ci);
      if (two) {
        out[2 * ((size_t)(i + 1) * h + k)] = 0.5 * (zi - ci);
        out[2 * ((size_t)(i + 1) * h + k) + 1] = -0.5 * (zr - cr);
      }
    }
  }
  free(z);
  double *col = dvec(2 * (size_t)n0);
#pragma omp for schedule(static)
  for (int k = 0; k < h; k++) {
    for (int i = 0; i < n0; i++) {
      col[2 * i] = out[2 * ((size_t)i * h + k)];
      col[2 * i + 1] = out[2 * ((size_t)i * h + k) + 1];
    }
    or_fft_c2c(n0, col, -1);
    for (int i = 0; i < n0; i++) {
      out[2 * ((size_t)i * h + k)] = col[2 * i];
      out[2 * ((size_t)i * h + k) + 1] = col[2 * i + 1];
    }
  }
  free(col);
  }
}

/* 2-D half-complex-to-real, FFTW c2r semantics: backward c2c along dimension 0
 * on every stored column, then a 1-D c2r along dimension 1 per row, which uses
 * only the real part of the DC (and, n1 even, the Nyquist) entry. */
void or_fft2_c2r(int n0, int n1, const double *in, double *out) {
  const int h = n1 / 2 + 1;
  double *t = dvec(2 * (size_t)n0 * h);
  memcpy(t, in, sizeof(double) * 2 * (size_t)n0 * h);
  (void)fft_plan(n0);
  (void)fft_plan(n1);
#pragma omp parallel
  {
  double *col = dvec(2 * (size_t)n0);
#pragma omp for schedule(static)
  for (int k = 0; k < h; k++) {
    for (int i = 0; i < n0; i++) {
      col[2 * i] = t[2 * ((size_t)i * h + k)];
      col[2 * i + 1] = t[2 * ((size_t)i * h + k) + 1];
    }
    or_fft_c2c(n0, col, +1);
    for (int i = 0; i < n0; i++) {
      t[2 * ((size_t)i * h + k)] = col[2 * i];
      t[2 * ((size_t)i * h + k) + 1] = col[2 * i + 1];
    }
  }
  free(col);
  double *z = dvec(2 * (size_t)n1);
#pragma omp for schedule(static)
  for (int i = 0; i < n0; i += 2) {
    const int two = (i + 1 < n0);
    const double *a = t + 2 * (size_t)i * h;
    const double *b = two ? t + 2 * (size_t)(i + 1) * h : NULL;
    for (int k = 0; k < h; k++) {
      double ar = a[2 * k], ai = a[2 * k + 1];
      double br = b ? b[2 * k] : 0.0, bi = b ? b[2 * k + 1] : 0.0;
      if (k == 0 || (2 * k == n1)) {
        ai = 0.0;
        bi = 0.0;
      }
      /* Z[k] = A[k] + i B[k] */
      z[2 * k] = ar - bi;
      z[2 * k + 1] = ai + br;
      int kk = n1 - k;
      if (k > 0 && kk > k && kk < n1) {
        /* Z[-k] = conj A[k] + i conj B[k] */
        z[2 * kk] = ar + bi;
        z[2 * kk + 1] = -ai + br;
      }
    }
    or_fft_c2c(n1, z, +1);
    for (int j = 0; j < n1; j++) {
      out[(size_t)i * n1 + j] = z[2 * j];
      if (two) out[(size_t)(i + 1) * n1 + j] = z[2 * j + 1];
    }
  }
  free(z);
  }
  free(t);
}

/* ======================================================================
 * Elementwise tools -- c/mfista_tools.c
 * ====================================================================== */

static double dasum(int n, const double *x) {
  double s = 0;
  for (int i = 0; i < n; i++) s += fabs(x[i]);
  return s;
}

/* c/mfista_tools.c:61-77: buf = x1 - x2; returns -<buf, AyAz> + c/2 <buf, buf> */
double or_calc_Q_part(int N, const double *x1, const double *x2, double c,
                      const double *AyAz, double *buf) {
  double t2 = 0, t3 = 0;
  for (int i = 0; i < N; i++) buf[i] = x1[i] - x2[i];
  for (int i = 0; i < N; i++) t2 += AyAz[i] * buf[i];
  for (int i = 0; i < N; i++) t3 += buf[i] * buf[i];
  return -t2 + c * t3 / 2;
}

/* c/mfista_tools.c:81-112 (soft_threshold, soft_threshold_box) */
void or_soft_threshold_box(const double *v, int n, double eta, double *out, int box_flag,
                           const float *cl_box) {
  for (int i = 0; i < n; i++) {
    double a = v[i];
    if (box_flag != 0 && cl_box[i] == 0)
      out[i] = 0;
    else if (a >= eta)
      out[i] = a - eta;
    else if (a <= -eta)
      out[i] = a + eta;
    else
      out[i] = 0;
  }
}

/* c/mfista_tools.c:114-141 (soft_threshold_nonneg, _nonneg_box) */
void or_soft_threshold_nonneg_box(const double *v, int n, double eta, double *out,
                                  int box_flag, const float *cl_box) {
  for (int i = 0; i < n; i++) {
    double a = v[i];
    if ((box_flag != 0 && cl_box[i] == 0) || a < eta)
      out[i] = 0;
    else
      out[i] = a - eta;
  }
}

/* c/mfista_tools.c:145-163.  Indexing x[NX*j+i], i < NX, j < NY. */
double or_TSV(int NX, int NY, const double *x) {
  double tsv = 0;
  for (int i = 0; i < NX - 1; ++i)
    for (int j = 0; j < NY - 1; ++j) {
      double a = x[NX * j + i] - x[NX * j + i + 1];
      double b = x[NX * j + i] - x[NX * (j + 1) + i];
      tsv += a * a;
      tsv += b * b;
    }
  for (int i = 0; i < NX - 1; ++i) {
    double a = x[NX * (NY - 1) + i] - x[NX * (NY - 1) + i + 1];
    tsv += a * a;
  }
  for (int j = 0; j < NY - 1; ++j) {
    double a = x[NX * j + (NX - 1)] - x[NX * (j + 1) + (NX - 1)];
    tsv += a * a;
  }
  return tsv;
}

/* c/mfista_tools.c:165-186: d = dTSV/dx = 2 sum over the 4-neighbours of (x_c - x_n) */
void or_d_TSV(int NX, int NY, const double *x, double *d) {
  for (int j = 0; j < NY; j++) d[NX * j + NX - 1] = 0;
  for (int i = 0; i < NX - 1; i++)
    for (int j = 0; j < NY; j++) d[NX * j + i] = 2 * (x[NX * j + i] - x[NX * j + i + 1]);
  for (int i = 1; i < NX; i++)
    for (int j = 0; j < NY; j++) d[NX * j + i] += 2 * (x[NX * j + i] - x[NX * j + i - 1]);
  for (int i = 0; i < NX; i++)
    for (int j = 0; j < NY - 1; j++) d[NX * j + i] += 2 * (x[NX * j + i] - x[NX * (j + 1) + i]);
  for (int i = 0; i < NX; i++)
    for (int j = 1; j < NY; j++) d[NX * j + i] += 2 * (x[NX * j + i] - x[NX * (j - 1) + i]);
}

/* ======================================================================
 * TV and the FGP prox -- c/mfista_TV_lib.c
 * ====================================================================== */

/* c/mfista_TV_lib.c:30-46 */
double or_TV(int NX, int NY, const double *x) {
  double tv = 0;
  for (int i = 0; i < NX - 1; ++i)
    for (int j = 0; j < NY - 1; ++j) {
      double a = x[NX * j + i] - x[NX * j + i + 1];
      double b = x[NX * j + i] - x[NX * (j + 1) + i];
      tv += sqrt(a * a + b * b);
    }
  for (int i = 0; i < NX - 1; ++i) tv += fabs(x[NX * (NY - 1) + i] - x[NX * (NY - 1) + i + 1]);
  for (int j = 0; j < NY - 1; ++j) tv += fabs(x[NX * j + NX - 1] - x[NX * (j + 1) + NX - 1]);
  return tv;
}

/* c/mfista_TV_lib.c:48-57: p[(NX-1)j+i] = x_{i,j}-x_{i+1,j}; q[NX j+i] = x_{i,j}-x_{i,j+1} */
void or_Lx(int NX, int NY, const double *x, double *p, double *q) {
  for (int i = 0; i < NX - 1; ++i)
    for (int j = 0; j < NY; ++j) p[(NX - 1) * j + i] = x[NX * j + i] - x[NX * j + i + 1];
  for (int i = 0; i < NX; ++i)
    for (int j = 0; j < NY - 1; ++j) q[NX * j + i] = x[NX * j + i] - x[NX * (j + 1) + i];
}

/* c/mfista_TV_lib.c:59-75: x = L^T (p, q) */
void or_Lpq(int NX, int NY, const double *p, const double *q, double *x) {
  for (int i = 0; i < NX * NY; ++i) x[i] = 0;
  for (int i = 0; i < NX - 1; ++i)
    for (int j = 0; j < NY; ++j) {
      x[NX * j + i] += p[(NX - 1) * j + i];
      x[NX * j + i + 1] -= p[(NX - 1) * j + i];
    }
  for (int i = 0; i < NX; ++i)
    for (int j = 0; j < NY - 1; ++j) {
      x[NX * j + i] += q[NX * j + i];
      x[NX * (j + 1) + i] -= q[NX * j + i];
    }
}

/* c/mfista_TV_lib.c:84-93 */
static void P_positive_box(int N, double *v, int box_flag, const float *cl_box) {
  for (int i = 0; i < N; ++i)
    if ((box_flag != 0 && cl_box[i] == 0) || v[i] < 0) v[i] = 0;
}

/* c/mfista_TV_lib.c:95-122 */
void or_P_pqiso(int NX, int NY, const double *p, const double *q, double *r, double *s) {
  for (int i = 0; i < NX - 1; ++i)
    for (int j = 0; j < NY - 1; ++j) {
      double a = p[(NX - 1) * j + i], b = q[NX * j + i];
      double t = sqrt(a * a + b * b);
      if (t < 1) t = 1;
      r[(NX - 1) * j + i] = a / t;
      s[NX * j + i] = b / t;
    }
  for (int i = 0; i < NX - 1; ++i) {
    double t = fabs(p[(NX - 1) * (NY - 1) + i]);
    if (t < 1) t = 1;
    r[(NX - 1) * (NY - 1) + i] = p[(NX - 1) * (NY - 1) + i] / t;
  }
  for (int j = 0; j < NY - 1; ++j) {
    double t = fabs(q[NX * j + NX - 1]);
    if (t < 1) t = 1;
    s[NX * j + NX - 1] = q[NX * j + NX - 1] / t;
  }
}

/* FGP of Beck & Teboulle for the TV prox.  mode 0: L1 prox in the primal
 * step (c/mfista_TV_lib.c:189-253, FGP_L1_box); mode 1: projection on x>=0
 * (c/mfista_TV_lib.c:320-384, FGP_nonneg_box). */
static void fgp_core(int mode, int NX, int NY, const double *b, double lambda_l1,
                     double lambda_tv, int ITER, double *x, int box_flag, const float *cl_box) {
  const int N = NX * NY, Np = (NX - 1) * NY, Nq = NX * (NY - 1);
  double *p = dvec(Np), *q = dvec(Nq), *r = dvec(Np), *s = dvec(Nq), *np_ = dvec(Np),
         *nq = dvec(Nq);
  const double a3 = 1 / (8 * lambda_tv);
  double t = 1, tnew;
  for (int it = 0; it < ITER; ++it) {
    tnew = (1 + sqrt(1 + 4 * t * t)) / 2;
    or_Lpq(NX, NY, r, s, x);
    for (int i = 0; i < N; i++) x[i] = x[i] * (-lambda_tv) + b[i];
    if (mode == 0)
      or_soft_threshold_box(x, N, lambda_l1, x, box_flag, cl_box);
    else
      P_positive_box(N, x, box_flag, cl_box);
    or_Lx(NX, NY, x, np_, nq);
    for (int i = 0; i < Np; i++) r[i] += a3 * np_[i];
    for (int i = 0; i < Nq; i++) s[i] += a3 * nq[i];
    or_P_pqiso(NX, NY, r, s, np_, nq);
    double beta1 = 1 + (t - 1) / tnew, beta2 = -(t - 1) / tnew;
    for (int i = 0; i < Np; i++) r[i] = np_[i] * beta1 + beta2 * p[i];
    for (int i = 0; i < Nq; i++) s[i] = nq[i] * beta1 + beta2 * q[i];
    t = tnew;
    memcpy(p, np_, sizeof(double) * Np);
    memcpy(q, nq, sizeof(double) * Nq);
  }
  or_Lpq(NX, NY, p, q, x);
  for (int i = 0; i < N; i++) x[i] = x[i] * (-lambda_tv) + b[i];
  if (mode == 0)
    or_soft_threshold_box(x, N, lambda_l1, x, box_flag, cl_box);
  else
    P_positive_box(N, x, box_flag, cl_box);
  free(p); free(q); free(r); free(s); free(np_); free(nq);
}

void or_FGP_L1_box(int NX, int NY, const double *b, double lambda_l1, double lambda_tv,
                   int ITER, double *x, int box_flag, const float *cl_box) {
  fgp_core(0, NX, NY, b, lambda_l1, lambda_tv, ITER, x, box_flag, cl_box);
}

void or_FGP_nonneg_box(int NX, int NY, const double *b, double lambda_tv, int ITER, double *x,
                       int box_flag, const float *cl_box) {
  fgp_core(1, NX, NY, b, 0.0, lambda_tv, ITER, x, box_flag, cl_box);
}

/* ======================================================================
 * NUFFT (Greengard & Lee 2004 fast Gaussian gridding)
 * ====================================================================== */

typedef struct or_nufft {
  int preset, M, Nx, Ny, MSP;
  double *E1, *E2x, *E2y, *E3x, *E3y, *E4; /* E2: [k][2 MSP] */
  int *mx, *my, *yneg;
  double *rgrid; /* 2Nx x 2Ny real */
  double *cgrid; /* 2Nx x (Ny+1) complex */
  double *mbuf;  /* C++ preset: padded spectrum buffer */
  double *priv;  /* OpenMP build: per-thread gridding buffers (npriv of them) */
  int npriv;
} or_nufft;

/* c/mfista_nufft_lib.c:32-40 */
static int idx_fftw(int m, int Mr) {
  if (m >= 0 && m < Mr / 2) return m;
  if (m >= -Mr / 2) return m + Mr;
  return 2 * Mr;
}

/* c/mfista_nufft_lib.c:42-48: centred pixel -> padded-grid index */
static int m2mr(int id, int N) { return (id < N / 2) ? id + 3 * N / 2 : id - N / 2; }

static void nufft_free(or_nufft *P) {
  free(P->E1); free(P->E2x); free(P->E2y); free(P->E3x); free(P->E3y); free(P->E4);
  free(P->mx); free(P->my); free(P->yneg); free(P->rgrid); free(P->cgrid); free(P->mbuf);
  free(P->priv);
  memset(P, 0, sizeof(*P));
}

/* preNUFFT: c/mfista_nufft_lib.c:50-121 (C-24) and c++/mfista_nufft_lib.cpp:85-184
 * (CXX-12).  Returns -1 if any tap leaves the (2Nx, 2Ny) grid. */
static int nufft_init(or_nufft *P, int preset, const double *u, const double *v, int M,
                      int Nx, int Ny) {
  memset(P, 0, sizeof(*P));
  P->preset = preset;
  P->M = M;
  P->Nx = Nx;
  P->Ny = Ny;
  const int MSP = (preset == OR_PRESET_CXX12) ? 6 : 12;
  P->MSP = MSP;
  const double pi = M_PI, Mrx = 2.0 * Nx, Mry = 2.0 * Ny;
  P->E1 = dvec(M);
  P->E2x = dvec((size_t)2 * MSP * M);
  P->E2y = dvec((size_t)2 * MSP * M);
  P->E3x = dvec(MSP + 1);
  P->E3y = dvec(MSP + 1);
  P->E4 = dvec((size_t)Nx * Ny);
  P->mx = (int *)calloc(M ? M : 1, sizeof(int));
  P->my = (int *)calloc(M ? M : 1, sizeof(int));
  P->yneg = (int *)calloc(M ? M : 1, sizeof(int));
  P->rgrid = dvec((size_t)4 * Nx * Ny);
  P->cgrid = dvec((size_t)2 * 2 * Nx * (Ny + 1));
  int bad = 0;
  if (preset == OR_PRESET_C24) {
    const double taux = 12 / ((double)Nx * Nx), tauy = 12 / ((double)Ny * Ny);
    const double coeff = pi / sqrt(taux * tauy);
    const double t3x = pow(pi / Mrx, 2.0) / taux, t3y = pow(pi / Mry, 2.0) / tauy;
    for (int l = 0; l <= MSP; ++l) {
      P->E3x[l] = exp(-t3x * (double)(l * l));
      P->E3y[l] = exp(-t3y * (double)(l * l));
    }
    for (int i = 0; i < Nx; ++i) {
      double ti = (double)(i - Nx / 2), tx = taux * ti * ti;
      for (int j = 0; j < Ny; ++j) {
        double tj = (double)(j - Nx / 2); /* the reference uses Nx/2 here (square images) */
        P->E4[(size_t)Ny * i + j] = coeff * exp(tx + tauy * tj * tj);
      }
    }
    for (int k = 0; k < M; ++k) {
      double tx = round(u[k] * Mrx / (2 * pi)), ty = round(v[k] * Mry / (2 * pi));
      P->mx[k] = (int)tx;
      P->my[k] = (int)ty;
      double xix = (2 * pi * tx) / Mrx, xiy = (2 * pi * ty) / Mry;
      P->E1[k] = exp(-pow(u[k] - xix, 2.0) / (4 * taux) - pow(v[k] - xiy, 2.0) / (4 * tauy));
      double ax = pi * (u[k] - xix) / (Mrx * taux), ay = pi * (v[k] - xiy) / (Mry * tauy);
      double pw = (double)(-MSP + 1);
      for (int j = 0; j < 2 * MSP; ++j) {
        P->E2x[(size_t)2 * MSP * k + j] = exp(pw * ax);
        P->E2y[(size_t)2 * MSP * k + j] = exp(pw * ay);
        pw += 1.0;
      }
    }
  } else {
    const double tau = ((double)MSP) / ((double)Nx * Nx);
    const double coeff = ((double)Nx * Ny) * pi / ((double)MSP);
    const double t3x = (pi * pi) / (Mrx * Mrx * tau), t3y = (pi * pi) / (Mry * Mry * tau);
    for (int k = 0; k < M; ++k) {
      double uk = u[k], vk = v[k];
      if (v[k] >= 0) {
        P->yneg[k] = 0;
      } else {
        uk = -uk;
        vk = -vk;
        P->yneg[k] = 1;
      }
      double tx = floor(uk * Mrx / (2.0 * pi)), ty = floor(vk * Mry / (2.0 * pi));
      P->mx[k] = (int)tx;
      P->my[k] = (int)ty;
      double xix = (2.0 * pi * tx) / Mrx, xiy = (2.0 * pi * ty) / Mry;
      P->E1[k] = exp(-((uk - xix) * (uk - xix)) / (4.0 * tau) -
                     ((vk - xiy) * (vk - xiy)) / (4.0 * tau));
      double ax = pi * (uk - xix) / (Mrx * tau), ay = pi * (vk - xiy) / (Mry * tau);
      for (int j = 0; j < 2 * MSP; ++j) {
        double l = (double)(j - MSP + 1);
        P->E2x[(size_t)2 * MSP * k + j] = exp(l * (ax - l * t3x));
        P->E2y[(size_t)2 * MSP * k + j] = exp(l * (ay - l * t3y));
      }
    }
    for (int i = 0; i < Nx; ++i) {
      double ti = (double)(i - Nx / 2), tx = tau * ti * ti;
      for (int j = 0; j < Ny; ++j) {
        double tj = (double)(j - Ny / 2);
        P->E4[(size_t)i * Ny + j] = coeff * exp(tx + tau * tj * tj);
      }
    }
    P->mbuf = dvec((size_t)2 * (2 * Nx + 2 * MSP) * (Ny + 2 * MSP + 1));
  }
  /* FOV: every tap m + l, l in [-MSP+1, MSP], must stay inside (-N, N)
   * (c++/mfista_nufft_lib.cpp:155-156 warns; c/mfista_nufft_lib.c:32-40 would
   * index out of range). */
  for (int k = 0; k < M; ++k) {
    if (P->mx[k] < -Nx + MSP || P->mx[k] > Nx - MSP - 1) bad++;
    if (P->my[k] < -Ny + MSP || P->my[k] > Ny - MSP - 1) bad++;
  }
  return bad ? -1 : 0;
}

/* Forward A.x: C-24 = NUFFT2d2, c/mfista_nufft_lib.c:187-243; CXX-12 = NUFFT2d2,
 * c++/mfista_nufft_lib.cpp:440-504.  Output assigned (not accumulated). */
static void nufft_fwd(or_nufft *P, const double *x, double *yr, double *yi) {
  const int Nx = P->Nx, Ny = P->Ny, MSP = P->MSP, M = P->M, Mrx = 2 * Nx, Mry = 2 * Ny,
            Mh = Ny + 1, S = 2 * MSP;
  const double MM = (double)Mrx * Mry;
  memset(P->rgrid, 0, sizeof(double) * (size_t)Mrx * Mry);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nx; ++i) {
    int ix = m2mr(i, Nx);
    for (int j = 0; j < Ny; ++j)
      P->rgrid[(size_t)ix * Mry + m2mr(j, Ny)] = x[(size_t)i * Ny + j] * P->E4[(size_t)i * Ny + j];
  }
  or_fft2_r2c(Mrx, Mry, P->rgrid, P->cgrid);
  const double *G = P->cgrid;
  if (P->preset == OR_PRESET_C24) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < M; ++k) {
      double fr = 0, fi = 0;
      for (int ly = -MSP + 1; ly <= MSP; ++ly) {
        double f_r = P->E1[k] * P->E2y[(size_t)S * k + ly + MSP - 1] * P->E3y[abs(ly)];
        int idy = idx_fftw(P->my[k] + ly, Mry);
        if (idy < Ny + 1) {
          for (int lx = -MSP + 1; lx <= MSP; ++lx) {
            int idx = idx_fftw(P->mx[k] + lx, Mrx);
            double t = f_r * P->E2x[(size_t)S * k + lx + MSP - 1] * P->E3x[abs(lx)] / MM;
            fr += G[2 * ((size_t)idx * Mh + idy)] * t;
            fi += -G[2 * ((size_t)idx * Mh + idy) + 1] * t;
          }
        } else {
          idy = idx_fftw(-(P->my[k] + ly), Mry);
          if (idy < Ny + 1)
            for (int lx = -MSP + 1; lx <= MSP; ++lx) {
              int idx = idx_fftw(-(P->mx[k] + lx), Mrx);
              double t = f_r * P->E2x[(size_t)S * k + lx + MSP - 1] * P->E3x[abs(lx)] / MM;
              fr += G[2 * ((size_t)idx * Mh + idy)] * t;
              fi += G[2 * ((size_t)idx * Mh + idy) + 1] * t;
            }
        }
      }
      yr[k] = fr;
      yi[k] = fi;
    }
    return;
  }
  /* CXX-12: spectrum copied into a buffer padded by MSP on every side, with
   * rows = ky + MSP (ky in [-MSP, Ny+MSP]) and cols = kx + Nx + MSP. */
  const int R = Ny + S + 1, C = 2 * Nx + S;
  double *hb = P->mbuf;
  memset(hb, 0, sizeof(double) * 2 * (size_t)R * C);
#define HB(r, c) (hb + 2 * ((size_t)(r) * C + (c)))
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nx; ++i)
    for (int r = 0; r < Mh; ++r) {
      HB(MSP + r, i + Nx + MSP)[0] = G[2 * ((size_t)i * Mh + r)];
      HB(MSP + r, i + Nx + MSP)[1] = G[2 * ((size_t)i * Mh + r) + 1];
      HB(MSP + r, i + MSP)[0] = G[2 * ((size_t)(i + Nx) * Mh + r)];
      HB(MSP + r, i + MSP)[1] = G[2 * ((size_t)(i + Nx) * Mh + r) + 1];
    }
  for (int i = 0; i < MSP; ++i) {
    for (int j = 0; j < Nx; ++j) {
      const double *s = HB(S - i, MSP + 2 * Nx - 1 - j);
      HB(i, MSP + 1 + j)[0] = s[0];
      HB(i, MSP + 1 + j)[1] = -s[1];
    }
    for (int j = 0; j < Nx - 1; ++j) {
      const double *s = HB(S - i, MSP + Nx - 1 - j);
      HB(i, MSP + Nx + 1 + j)[0] = s[0];
      HB(i, MSP + Nx + 1 + j)[1] = -s[1];
    }
  }
  const double MMinv = 1 / (4.0 * (double)Nx * Ny);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < M; ++k) {
    double sr = 0, si = 0;
    for (int a = 0; a < S; a++) {
      double gy = P->E2y[(size_t)S * k + a];
      for (int b = 0; b < S; b++) {
        double g = gy * P->E2x[(size_t)S * k + b];
        const double *h = HB(P->my[k] + 1 + a, P->mx[k] + Nx + 1 + b);
        sr += g * h[0];
        si += g * h[1];
      }
    }
    double fr = MMinv * P->E1[k] * sr, fi = -MMinv * P->E1[k] * si; /* conj */
    if (P->yneg[k]) fi = -fi;
    yr[k] = fr;
    yi[k] = fi;
  }
#undef HB
}

/* Adjoint Re(A^H r): C-24 = NUFFT2d1, c/mfista_nufft_lib.c:123-185; CXX-12 =
 * NUFFT2d1_sngl, c++/mfista_nufft_lib.cpp:189-258 (the tiled NUFFT2d1 :260-438
 * computes the same sums). */
static void nufft_adj(or_nufft *P, const double *Fr, const double *Fi, double *xo) {
  const int Nx = P->Nx, Ny = P->Ny, MSP = P->MSP, M = P->M, Mrx = 2 * Nx, Mry = 2 * Ny,
            Mh = Ny + 1, S = 2 * MSP;
  const double MM = (double)Mrx * Mry;
  double *H = P->cgrid;
  memset(H, 0, sizeof(double) * 2 * (size_t)Mrx * Mh);
  if (P->preset == OR_PRESET_C24) {
    for (int k = 0; k < M; ++k) {
      double v0r = Fr[k] * P->E1[k], v0i = -Fi[k] * P->E1[k];
      for (int ly = -MSP + 1; ly <= MSP; ++ly) {
        double t = P->E2y[(size_t)S * k + ly + MSP - 1] * P->E3y[abs(ly)];
        double vyr = v0r * t, vyi = v0i * t;
        int idy = idx_fftw(P->my[k] + ly, Mry);
        if (idy < Ny + 1)
          for (int lx = -MSP + 1; lx <= MSP; ++lx) {
            int idx = idx_fftw(P->mx[k] + lx, Mrx);
            double w = P->E2x[(size_t)S * k + lx + MSP - 1] * P->E3x[abs(lx)] / 2;
            H[2 * ((size_t)idx * Mh + idy)] += vyr * w;
            H[2 * ((size_t)idx * Mh + idy) + 1] += vyi * w;
          }
        idy = idx_fftw(-(P->my[k] + ly), Mry);
        if (idy < Ny + 1)
          for (int lx = -MSP + 1; lx <= MSP; ++lx) {
            int idx = idx_fftw(-(P->mx[k] + lx), Mrx);
            double w = P->E2x[(size_t)S * k + lx + MSP - 1] * P->E3x[abs(lx)] / 2;
            H[2 * ((size_t)idx * Mh + idy)] += vyr * w;
            H[2 * ((size_t)idx * Mh + idy) + 1] -= vyi * w;
          }
      }
    }
    or_fft2_c2r(Mrx, Mry, H, P->rgrid);
    for (int k = 0; k < Nx; ++k) {
      int idx = m2mr(k, Nx);
      for (int j = 0; j < Ny; ++j)
        xo[(size_t)k * Ny + j] =
            P->rgrid[(size_t)idx * Mry + m2mr(j, Ny)] * P->E4[(size_t)k * Ny + j] / MM;
    }
    return;
  }
  /* CXX-12: padded buffer rows = kx + Nx + MSP, cols = ky + MSP */
  const int R = 2 * Nx + S, C = Ny + S + 1;
  double *lb = P->mbuf;
  memset(lb, 0, sizeof(double) * 2 * (size_t)R * C);
#define LB(r, c) (lb + 2 * ((size_t)(r) * C + (c)))
  const double MMinv = 1 / (2.0 * (4.0 * (double)Nx * Ny));
#define SCATTER(buf, k)                                                       \
  do {                                                                        \
    double vr = Fr[k], vi = -Fi[k]; /* NU_SIGN = -1: conj */                  \
    if (P->yneg[k]) vi = -vi;                                                 \
    double v0r = MMinv * P->E1[k] * vr, v0i = MMinv * P->E1[k] * vi;          \
    for (int a = 0; a < S; a++) {                                             \
      double gx = P->E2x[(size_t)S * (k) + a];                                \
      for (int b = 0; b < S; b++) {                                           \
        double g = gx * P->E2y[(size_t)S * (k) + b];                          \
        double *d = (buf) + 2 * ((size_t)(P->mx[k] + 1 + Nx + a) * C + (P->my[k] + 1 + b)); \
        d[0] += v0r * g;                                                      \
        d[1] += v0i * g;                                                      \
      }                                                                       \
    }                                                                         \
  } while (0)
#ifdef _OPENMP
  const int nthr = omp_get_max_threads();
  if (nthr > 1) { /* one grid per thread, summed in thread order afterwards */
    const size_t nb = 2 * (size_t)R * C;
    if (!P->priv || P->npriv < nthr) {
      free(P->priv);
      P->priv = dvec(nb * (size_t)nthr);
      P->npriv = nthr;
    }
#pragma omp parallel num_threads(nthr)
    {
      double *mine = P->priv + nb * (size_t)omp_get_thread_num();
      memset(mine, 0, sizeof(double) * nb);
#pragma omp for schedule(static)
      for (int k = 0; k < M; ++k) SCATTER(mine, k);
#pragma omp for schedule(static)
      for (size_t e = 0; e < nb; e++) {
        double acc = 0;
        for (int q = 0; q < nthr; q++) acc += P->priv[nb * (size_t)q + e];
        lb[e] = acc;
      }
    }
  } else
#endif
  {
    for (int k = 0; k < M; ++k) SCATTER(lb, k);
  }
#undef SCATTER
  /* fold the conjugate (ky < 0) half onto ky > 0 */
  for (int i = 0; i < Nx; ++i)
    for (int j = 0; j < MSP; ++j) {
      double *d = LB(MSP + Nx + i, MSP + 1 + j);
      const double *s = LB(MSP + Nx - i, MSP - 1 - j);
      d[0] += s[0];
      d[1] -= s[1];
    }
  for (int i = 0; i < Nx - 1; ++i)
    for (int j = 0; j < MSP; ++j) {
      double *d = LB(MSP + 1 + i, MSP + 1 + j);
      const double *s = LB(MSP + 2 * Nx - 1 - i, MSP - 1 - j);
      d[0] += s[0];
      d[1] -= s[1];
    }
  for (int i = 0; i < MSP - 1; ++i) {
    double *d = LB(MSP, MSP + 1 + i);
    const double *s = LB(MSP, MSP - 1 - i);
    d[0] += s[0];
    d[1] -= s[1];
  }
  /* symmetrise the ky = 0 and ky = Ny lines */
  double *ax = dvec(4 * (size_t)Nx);
  for (int line = 0; line < 2; line++) {
    int col = MSP + (line ? Ny : 0);
    for (int i = 0; i < 2 * Nx; ++i) {
      ax[2 * i] = LB(MSP + i, col)[0];
      ax[2 * i + 1] = LB(MSP + i, col)[1];
    }
    for (int i = 0; i < Nx - 1; ++i) {
      LB(MSP + Nx + 1 + i, col)[0] += ax[2 * (Nx - 1 - i)];
      LB(MSP + Nx + 1 + i, col)[1] -= ax[2 * (Nx - 1 - i) + 1];
    }
    for (int i = 0; i < Nx - 1; ++i) {
      LB(MSP + 1 + i, col)[0] += ax[2 * (2 * Nx - 1 - i)];
      LB(MSP + 1 + i, col)[1] -= ax[2 * (2 * Nx - 1 - i) + 1];
    }
  }
  free(ax);
  LB(MSP + Nx, MSP)[0] *= 2.0;
  LB(MSP + Nx, MSP)[1] *= 2.0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < Nx; ++i)
    for (int r = 0; r < Mh; ++r) {
      H[2 * ((size_t)i * Mh + r)] = LB(i + MSP + Nx, MSP + r)[0];
      H[2 * ((size_t)i * Mh + r) + 1] = LB(i + MSP + Nx, MSP + r)[1];
      H[2 * ((size_t)(i + Nx) * Mh + r)] = LB(i + MSP, MSP + r)[0];
      H[2 * ((size_t)(i + Nx) * Mh + r) + 1] = LB(i + MSP, MSP + r)[1];
    }
#undef LB
  or_fft2_c2r(Mrx, Mry, H, P->rgrid);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < Nx; ++k) {
    int idx = m2mr(k, Nx);
    for (int j = 0; j < Ny; ++j)
      xo[(size_t)k * Ny + j] = P->rgrid[(size_t)idx * Mry + m2mr(j, Ny)] * P->E4[(size_t)k * Ny + j];
  }
}

int or_nufft_forward(int preset, const double *u, const double *v, int M, int Nx, int Ny,
                     const double *x, double *yr, double *yi) {
  or_nufft P;
  int rc = nufft_init(&P, preset, u, v, M, Nx, Ny);
  if (rc == 0) nufft_fwd(&P, x, yr, yi);
  nufft_free(&P);
  return rc;
}

int or_nufft_adjoint(int preset, const double *u, const double *v, int M, int Nx, int Ny,
                     const double *yr, const double *yi, double *x) {
  or_nufft P;
  int rc = nufft_init(&P, preset, u, v, M, Nx, Ny);
  if (rc == 0) nufft_adj(&P, yr, yi, x);
  nufft_free(&P);
  return rc;
}

/* ======================================================================
 * Data-term operators: F(x) = chi^2/2 and g = -dF/dx at the last F point
 * ====================================================================== */

typedef struct or_op {
  double (*F)(void *ctx, const double *x);
  void (*grad)(void *ctx, double *g);
  void *ctx;
} or_op;

typedef struct {
  or_nufft P;
  const double *vr, *vi;
  double *w, *yr, *yi;
} nufft_ctx;

/* calc_F_part_nufft: c/mfista_nufft_lib.c:245-285; c++/mfista_nufft_lib.cpp:506-531 */
static double nufft_F(void *vctx, const double *x) {
  nufft_ctx *C = (nufft_ctx *)vctx;
  const int M = C->P.M;
  nufft_fwd(&C->P, x, C->yr, C->yi);
  double sr = 0, si = 0;
  for (int i = 0; i < M; ++i) {
    C->yr[i] = (C->vr[i] - C->yr[i]) * C->w[i];
    C->yi[i] = (C->vi[i] - C->yi[i]) * C->w[i];
  }
  for (int i = 0; i < M; ++i) sr += C->yr[i] * C->yr[i];
  for (int i = 0; i < M; ++i) si += C->yi[i] * C->yi[i];
  return (sr + si) / 2;
}

/* dF_dx_nufft: c/mfista_nufft_lib.c:287-304; c++/mfista_nufft_lib.cpp:533-546 */
static void nufft_grad(void *vctx, double *g) {
  nufft_ctx *C = (nufft_ctx *)vctx;
  for (int i = 0; i < C->P.M; ++i) {
    C->yr[i] *= C->w[i];
    C->yi[i] *= C->w[i];
  }
  nufft_adj(&C->P, C->yr, C->yi, g);
}

typedef struct {
  int NX, NY, NYh;
  double *yf_h, *mask_h, *x4f, *r_h, *full;
} fft_ctx;

/* idx2mat + full2half/fft_full2half: c/mfista_fft_lib.c:27-114 */
static void fft_ctx_init(fft_ctx *C, int M, int NX, int NY, const int *u_idx, const int *v_idx,
                         const double *y_r, const double *y_i, const double *noise) {
  C->NX = NX;
  C->NY = NY;
  C->NYh = NY / 2 + 1;
  double *mk = dvec((size_t)NX * NY), *yf = dvec(2 * (size_t)NX * NY);
  for (int i = 0; i < M; ++i) {
    size_t id = (size_t)NY * u_idx[i] + v_idx[i];
    mk[id] = 1 / noise[i];
    yf[2 * id] = y_r[i] / noise[i];
    yf[2 * id + 1] = y_i[i] / noise[i];
  }
  C->mask_h = dvec((size_t)NX * C->NYh);
  C->yf_h = dvec(2 * (size_t)NX * C->NYh);
  for (int i = 0; i < NX; ++i)
    for (int j = 0; j < C->NYh; ++j) {
      C->mask_h[(size_t)C->NYh * i + j] = mk[(size_t)NY * i + j];
      C->yf_h[2 * ((size_t)C->NYh * i + j)] = yf[2 * ((size_t)NY * i + j)];
      C->yf_h[2 * ((size_t)C->NYh * i + j) + 1] = yf[2 * ((size_t)NY * i + j) + 1];
    }
  free(mk);
  free(yf);
  C->x4f = dvec((size_t)NX * NY);
  C->r_h = dvec(2 * (size_t)NX * C->NYh);
  C->full = dvec(2 * (size_t)NX * NY);
}

static void fft_ctx_free(fft_ctx *C) {
  free(C->yf_h); free(C->mask_h); free(C->x4f); free(C->r_h); free(C->full);
}

/* fft_half2full: c/mfista_fft_lib.c:27-48 (even NX, NY) */
static void half2full(int NX, int NY, const double *h, double *f) {
  const int NXh = NX / 2 + 1, NYh = NY / 2 + 1;
#define F_(i, j) (f + 2 * ((size_t)NY * (i) + (j)))
#define H_(i, j) (h + 2 * ((size_t)NYh * (i) + (j)))
  for (int i = 0; i < NX; ++i)
    for (int j = 0; j < NYh; ++j) {
      F_(i, j)[0] = H_(i, j)[0];
      F_(i, j)[1] = H_(i, j)[1];
    }
  for (int i = 1; i < NXh - 1; ++i)
    for (int j = 1; j < NYh - 1; ++j) {
      F_(NX - i, NY - j)[0] = H_(i, j)[0];
      F_(NX - i, NY - j)[1] = -H_(i, j)[1];
    }
  for (int i = NXh; i < NX; ++i)
    for (int j = 1; j < NYh - 1; ++j) {
      F_(NX - i, NY - j)[0] = H_(i, j)[0];
      F_(NX - i, NY - j)[1] = -H_(i, j)[1];
    }
  for (int j = 1; j < NYh - 1; ++j) {
    F_(0, NY - j)[0] = H_(0, j)[0];
    F_(0, NY - j)[1] = -H_(0, j)[1];
    F_(NXh - 1, NY - j)[0] = H_(NXh - 1, j)[0];
    F_(NXh - 1, NY - j)[1] = -H_(NXh - 1, j)[1];
  }
#undef F_
#undef H_
}

/* calc_F_part_fft + calc_yAx_fft + norm_fftw_complex: c/mfista_fft_lib.c:116-159 */
static double fft_F(void *vctx, const double *x) {
  fft_ctx *C = (fft_ctx *)vctx;
  const int NX = C->NX, NY = C->NY, NYh = C->NYh;
  const double sq = sqrt((double)NX * NY);
  memcpy(C->x4f, x, sizeof(double) * (size_t)NX * NY);
  or_fft2_r2c(NX, NY, C->x4f, C->r_h);
  for (size_t i = 0; i < (size_t)NX * NYh; ++i) {
    C->r_h[2 * i] = C->yf_h[2 * i] - C->mask_h[i] * C->r_h[2 * i] / sq;
    C->r_h[2 * i + 1] = C->yf_h[2 * i + 1] - C->mask_h[i] * C->r_h[2 * i + 1] / sq;
  }
  half2full(NX, NY, C->r_h, C->full);
  double s = 0;
  for (size_t i = 0; i < (size_t)NX * NY; ++i)
    s += C->full[2 * i] * C->full[2 * i] + C->full[2 * i + 1] * C->full[2 * i + 1];
  return s / 4;
}

/* dF_dx_fft: c/mfista_fft_lib.c:161-175 */
static void fft_grad(void *vctx, double *g) {
  fft_ctx *C = (fft_ctx *)vctx;
  const int NX = C->NX, NY = C->NY;
  const double sq = sqrt((double)NX * NY);
  for (size_t i = 0; i < (size_t)NX * C->NYh; ++i) {
    C->r_h[2 * i] *= C->mask_h[i] / (2 * sq);
    C->r_h[2 * i + 1] *= C->mask_h[i] / (2 * sq);
  }
  or_fft2_c2r(NX, NY, C->r_h, g);
}

double or_fft_F_grad(int M, int NX, int NY, const int *u_idx, const int *v_idx,
                     const double *y_r, const double *y_i, const double *noise,
                     const double *x, double *g) {
  fft_ctx C;
  fft_ctx_init(&C, M, NX, NY, u_idx, v_idx, y_r, y_i, noise);
  double F = fft_F(&C, x);
  if (g) fft_grad(&C, g);
  fft_ctx_free(&C);
  return F;
}

typedef struct {
  int M, N;
  const double *y, *A;
  double *r;
} dft_ctx;

/* calc_yAz + calc_F_part: c/mfista_dft_lib.c:30-66 (A column-major M x N) */
static double dft_F(void *vctx, const double *x) {
  dft_ctx *C = (dft_ctx *)vctx;
  for (int i = 0; i < C->M; i++) C->r[i] = C->y[i];
  for (int j = 0; j < C->N; j++) {
    const double xj = x[j];
    const double *a = C->A + (size_t)C->M * j;
    for (int i = 0; i < C->M; i++) C->r[i] -= a[i] * xj;
  }
  double s = 0;
  for (int i = 0; i < C->M; i++) s += C->r[i] * C->r[i];
  return s / 2;
}

/* dL_dx: c/mfista_dft_lib.c:41-49, g = A^T (y - A z) */
static void dft_grad(void *vctx, double *g) {
  dft_ctx *C = (dft_ctx *)vctx;
  for (int j = 0; j < C->N; j++) {
    const double *a = C->A + (size_t)C->M * j;
    double s = 0;
    for (int i = 0; i < C->M; i++) s += a[i] * C->r[i];
    g[j] = s;
  }
}

/* ======================================================================
 * The MFISTA loop (Appendix: Beck & Teboulle monotone FISTA with
 * backtracking).  One restatement serves the three modes, whose loops are
 * identical apart from the operator:
 *   TSV: c/mfista_nufft_lib.c:557-787, c/mfista_fft_lib.c:406-609,
 *        c/mfista_dft_lib.c:282-436, c++/mfista_nufft_lib.cpp:680-927
 *   TV:  c/mfista_nufft_lib.c:308-553, c/mfista_fft_lib.c:179-402,
 *        c/mfista_dft_lib.c:70-249
 * eta0 = backtracking factor of the first outer iteration (10 for the C++
 * NUFFT, c++/mfista_nufft_lib.cpp:818,860-865; 1.1 otherwise).
 * Returns ITER; x holds the final image; *cinit receives the final c.
 * ====================================================================== */
static int mfista_loop(const or_op *op, int tv_mode, int NX, int NY, int maxiter, double eps,
                       double lambda_l1, double lambda_tv, double lambda_tsv, double *cinit,
                       double *x, int nonneg_flag, int box_flag, const float *cl_box,
                       double eta0, or_opts *opts) {
  const int NN = NX * NY;
  const int fgp_iter = (opts && opts->fgp_iter > 0) ? opts->fgp_iter : OR_FGPITER;
  const double rel_tol = opts ? opts->rel_tol : 0.0;
  const int verbose = opts ? opts->verbose : 0;
  double *cost = dvec(maxiter > 0 ? maxiter : 1);
  double *z = dvec(NN), *xnew = dvec(NN), *xtmp = dvec(NN), *dfdx = dvec(NN), *dtmp = dvec(NN);
  memcpy(z, x, sizeof(double) * NN);
  double c = *cinit, mu = 1, munew, Qcore, Fval = 0, Qval, costtmp;
  double eta = eta0;
  int iter, max_trials = 0;

  costtmp = op->F(op->ctx, x) + lambda_l1 * dasum(NN, x);
  if (tv_mode)
    costtmp += lambda_tv * or_TV(NX, NY, x);
  else if (lambda_tsv > 0)
    costtmp += lambda_tsv * or_TSV(NX, NY, x);

  for (iter = 0; iter < maxiter; iter++) {
    cost[iter] = costtmp;
    if (verbose && (iter % 10) == 0) printf("%d cost = %f, c = %f \n", iter + 1, cost[iter], c);
    Qcore = op->F(op->ctx, z);
    op->grad(op->ctx, dfdx);
    if (!tv_mode && lambda_tsv > 0.0) {
      Qcore += lambda_tsv * or_TSV(NX, NY, z);
      or_d_TSV(NX, NY, z, dtmp);
      for (int k = 0; k < NN; k++) dfdx[k] += -lambda_tsv * dtmp[k];
    }
    int i;
    for (i = 0; i < maxiter; i++) {
      const double cinv = 1 / c;
      if (tv_mode && nonneg_flag == 1) {
        for (int k = 0; k < NN; k++) xtmp[k] = -lambda_l1 / c + cinv * dfdx[k] + z[k];
        or_FGP_nonneg_box(NX, NY, xtmp, lambda_tv / c, fgp_iter, xnew, box_flag, cl_box);
      } else {
        for (int k = 0; k < NN; k++) xtmp[k] = dfdx[k] * cinv + z[k];
        if (tv_mode)
          or_FGP_L1_box(NX, NY, xtmp, lambda_l1 / c, lambda_tv / c, fgp_iter, xnew, box_flag,
                        cl_box);
        else if (nonneg_flag == 1)
          or_soft_threshold_nonneg_box(xtmp, NN, lambda_l1 / c, xnew, box_flag, cl_box);
        else
          or_soft_threshold_box(xtmp, NN, lambda_l1 / c, xnew, box_flag, cl_box);
      }
      Fval = op->F(op->ctx, xnew);
      if (!tv_mode && lambda_tsv > 0.0) Fval += lambda_tsv * or_TSV(NX, NY, xnew);
      Qval = or_calc_Q_part(NN, xnew, z, c, dfdx, xtmp) + Qcore;
      if (Fval <= Qval) break;
      c *= eta;
    }
    if (i + 1 > max_trials) max_trials = i + 1;
    eta = OR_ETA;
    c /= eta;
    munew = (1 + sqrt(1 + 4 * mu * mu)) / 2;
    Fval += lambda_l1 * dasum(NN, xnew);
    if (tv_mode) Fval += lambda_tv * or_TV(NX, NY, xnew);
    if (Fval < cost[iter]) {
      costtmp = Fval;
      const double a = (1 - mu) / munew, b = 1 + ((mu - 1) / munew);
      for (int k = 0; k < NN; k++) z[k] = x[k] * a + b * xnew[k];
      memcpy(x, xnew, sizeof(double) * NN);
    } else {
      const double a = 1 - (mu / munew), b = mu / munew;
      for (int k = 0; k < NN; k++) z[k] = x[k] * a + b * xnew[k];
      if ((iter > 1) && (dasum(NN, x) == 0)) break;
    }
    if (iter >= OR_MINITER) {
      double d = cost[iter - OR_TD] - cost[iter];
      if (rel_tol > 0 ? (d <= rel_tol * cost[iter]) : (d < eps)) break;
    }
    mu = munew;
  }
  if (iter == maxiter) iter = iter - 1;
  if (opts) {
    opts->max_trials_out = max_trials;
    if (opts->cost_trace) memcpy(opts->cost_trace, cost, sizeof(double) * (iter + 1));
  }
  *cinit = c;
  free(cost); free(z); free(xnew); free(xtmp); free(dfdx); free(dtmp);
  return iter + 1;
}

static double now_sec(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* l1 cost, N_active and the regulariser terms shared by every calc_result_* */
static void result_common(or_result *R, int NX, int NY, double lambda_l1, double lambda_tv,
                          double lambda_tsv, const double *x) {
  const int NN = NX * NY;
  R->N = NN;
  R->NX = NX;
  R->NY = NY;
  R->lambda_l1 = lambda_l1;
  R->lambda_tv = lambda_tv;
  R->lambda_tsv = lambda_tsv;
  R->l1cost = 0;
  R->N_active = 0;
  for (int i = 0; i < NN; ++i) {
    double t = fabs(x[i]);
    if (t > 0) {
      R->l1cost += t;
      ++R->N_active;
    }
  }
  R->finalcost = R->sq_error / 2;
  if (lambda_l1 > 0) R->finalcost += lambda_l1 * R->l1cost;
  if (lambda_tsv > 0) {
    R->tsvcost = or_TSV(NX, NY, x);
    R->finalcost += lambda_tsv * R->tsvcost;
  } else if (lambda_tv > 0) {
    R->tvcost = or_TV(NX, NY, x);
    R->finalcost += lambda_tv * R->tvcost;
  }
}

/* calc_result_nufft: c/mfista_nufft_lib.c:791-905 (C-24: RESULT.M = M/2);
 * calc_costs_nufft: c++/mfista_nufft_lib.cpp:931-996 (CXX-12: RESULT.M = M) */
int or_result_nufft(int preset, const double *u, const double *v, const double *vis_r,
                    const double *vis_i, const double *vis_std, int M, int NX, int NY,
                    double lambda_l1, double lambda_tv, double lambda_tsv, const double *x,
                    or_result *R) {
  nufft_ctx C;
  if (nufft_init(&C.P, preset, u, v, M, NX, NY) != 0) {
    nufft_free(&C.P);
    return -1;
  }
  C.vr = vis_r;
  C.vi = vis_i;
  C.w = dvec(M);
  C.yr = dvec(M);
  C.yi = dvec(M);
  for (int i = 0; i < M; ++i) C.w[i] = 1 / vis_std[i];
  double t = nufft_F(&C, x);
  R->sq_error = 2 * t;
  R->M = (preset == OR_PRESET_CXX12) ? M : (int)(M / 2);
  R->mean_sq_error = R->sq_error / ((double)M);
  result_common(R, NX, NY, lambda_l1, lambda_tv, lambda_tsv, x);
  nufft_free(&C.P);
  free(C.w); free(C.yr); free(C.yi);
  return 0;
}

/* mfista_imaging_core_nufft: c/mfista_nufft_lib.c:909-957, c++/mfista_nufft_lib.cpp:1016-1074.
 * Returns 0; -1 both lambda_tv and lambda_tsv positive (RESULT untouched);
 * -2 TV requested with the C++ preset (not implemented there); -3 FOV violation. */
int or_imaging_core_nufft(const double *u_dx, const double *v_dy, const double *vis_r,
                          const double *vis_i, const double *vis_std, int M, int NX, int NY,
                          int maxiter, double eps, double lambda_l1, double lambda_tv,
                          double lambda_tsv, double cinit, const double *xinit, double *xout,
                          int nonneg_flag, int box_flag, const float *cl_box, or_result *R,
                          or_opts *opts) {
  const int preset = opts ? opts->preset : OR_PRESET_C24;
  double epsilon = 0, c = cinit;
  for (int i = 0; i < M; ++i) epsilon += vis_r[i] * vis_r[i];
  for (int i = 0; i < M; ++i) epsilon += vis_i[i] * vis_i[i];
  epsilon *= eps / ((double)M);
  int tv_mode;
  if (lambda_tv == 0)
    tv_mode = 0;
  else if (lambda_tsv == 0)
    tv_mode = 1;
  else {
    printf("You cannot set both of lambda_TV and lambda_TSV positive.\n");
    return -1;
  }
  if (tv_mode && preset == OR_PRESET_CXX12) return -2;
  const int NN = NX * NY;
  double t0 = now_sec();
  nufft_ctx C;
  if (nufft_init(&C.P, preset, u_dx, v_dy, M, NX, NY) != 0) {
    nufft_free(&C.P);
    printf("Error: the field of view is too small for the (u,v) coverage.\n");
    return -3;
  }
  C.vr = vis_r;
  C.vi = vis_i;
  C.w = dvec(M);
  C.yr = dvec(M);
  C.yi = dvec(M);
  for (int i = 0; i < M; ++i) C.w[i] = 1 / vis_std[i];
  or_op op = {nufft_F, nufft_grad, &C};
  int iter = 0;
  if (!tv_mode && nonneg_flag != 0 && nonneg_flag != 1) {
    printf("nonneg_flag must be chosen properly.\n");
    iter = 0;
  } else {
    memcpy(xout, xinit, sizeof(double) * NN);
    iter = mfista_loop(&op, tv_mode, NX, NY, maxiter, epsilon, lambda_l1, lambda_tv, lambda_tsv,
                       &c, xout, nonneg_flag, box_flag, cl_box,
                       preset == OR_PRESET_CXX12 ? 10.0 : OR_ETA, opts);
  }
  double t1 = now_sec();
  nufft_free(&C.P);
  free(C.w); free(C.yr); free(C.yi);
  R->comp_time = t1 - t0;
  R->ITER = iter;
  R->nonneg = nonneg_flag;
  R->Lip_const = c;
  R->maxiter = maxiter;
  or_result_nufft(preset, u_dx, v_dy, vis_r, vis_i, vis_std, M, NX, NY, lambda_l1, lambda_tv,
                  lambda_tsv, xout, R);
  return 0;
}

/* mfista_imaging_core_fft: c/mfista_fft_lib.c:701-756 (+ calc_result_fft :613-697).
 * Returns 0; -1 both regularisers positive; -4 odd image size (the reference's
 * half2full reads unset entries then). */
int or_imaging_core_fft(const int *u_idx, const int *v_idx, const double *y_r,
                        const double *y_i, const double *noise_stdev, int M, int NX, int NY,
                        int maxiter, double eps, double lambda_l1, double lambda_tv,
                        double lambda_tsv, double cinit, const double *xinit, double *xout,
                        int nonneg_flag, int box_flag, const float *cl_box, or_result *R,
                        or_opts *opts) {
  double epsilon = 0, c = cinit;
  for (int i = 0; i < M; ++i) epsilon += y_r[i] * y_r[i] + y_i[i] * y_i[i];
  epsilon *= eps / ((double)M);
  if ((NX % 2) || (NY % 2)) return -4;
  int tv_mode;
  if (lambda_tv == 0)
    tv_mode = 0;
  else if (lambda_tsv == 0)
    tv_mode = 1;
  else {
    printf("You cannot set both of lambda_TV and lambda_TSV positive.\n");
    return -1;
  }
  const int NN = NX * NY;
  fft_ctx C;
  fft_ctx_init(&C, M, NX, NY, u_idx, v_idx, y_r, y_i, noise_stdev);
  or_op op = {fft_F, fft_grad, &C};
  double t0 = now_sec();
  int iter = 0;
  if (!tv_mode && nonneg_flag != 0 && nonneg_flag != 1) {
    printf("nonneg_flag must be chosen properly.\n");
  } else {
    memcpy(xout, xinit, sizeof(double) * NN);
    iter = mfista_loop(&op, tv_mode, NX, NY, maxiter, epsilon, lambda_l1, lambda_tv, lambda_tsv,
                       &c, xout, nonneg_flag, box_flag, cl_box, OR_ETA, opts);
  }
  double t1 = now_sec();
  R->comp_time = t1 - t0;
  R->ITER = iter;
  R->nonneg = nonneg_flag;
  R->Lip_const = c;
  R->maxiter = maxiter;
  double t = fft_F(&C, xout);
  R->sq_error = 2 * t;
  R->M = (int)(M / 2);
  R->mean_sq_error = R->sq_error / ((double)M);
  result_common(R, NX, NY, lambda_l1, lambda_tv, lambda_tsv, xout);
  fft_ctx_free(&C);
  return 0;
}

/* mfista_imaging_core_dft: c/mfista_dft_lib.c:540-589 (+ calc_result :440-536).
 * eps is used unscaled (c/mfista_dft_lib.c:559).  R->residual is malloc'ed
 * (M doubles); release it with or_free. */
int or_imaging_core_dft(const double *y, const double *A, int M, int N, int NX, int NY,
                        int maxiter, double eps, double lambda_l1, double lambda_tv,
                        double lambda_tsv, double cinit, const double *xinit, double *xout,
                        int nonneg_flag, int box_flag, const float *cl_box, or_result *R,
                        or_opts *opts) {
  double c = cinit;
  memcpy(xout, xinit, sizeof(double) * N);
  int tv_mode;
  if (lambda_tv == 0)
    tv_mode = 0;
  else if (lambda_tsv == 0)
    tv_mode = 1;
  else {
    printf("You cannot set both of lambda_TV and lambda_TSV positive.\n");
    return -1;
  }
  dft_ctx C = {M, N, y, A, dvec(M)};
  or_op op = {dft_F, dft_grad, &C};
  double t0 = now_sec();
  int iter = 0;
  if (!tv_mode && nonneg_flag != 0 && nonneg_flag != 1)
    printf("nonneg_flag must be chosen properly.\n");
  else
    iter = mfista_loop(&op, tv_mode, NX, NY, maxiter, eps, lambda_l1, lambda_tv, lambda_tsv, &c,
                       xout, nonneg_flag, box_flag, cl_box, OR_ETA, opts);
  double t1 = now_sec();
  R->comp_time = t1 - t0;
  R->ITER = iter;
  R->nonneg = nonneg_flag;
  R->Lip_const = c;
  R->maxiter = maxiter;
  dft_F(&C, xout);
  R->residual = dvec(M);
  R->sq_error = 0;
  for (int i = 0; i < M; i++) {
    R->sq_error += C.r[i] * C.r[i];
    R->residual[i] = C.r[i];
  }
  R->M = M;
  R->mean_sq_error = R->sq_error / ((double)M);
  result_common(R, NX, NY, lambda_l1, lambda_tv, lambda_tsv, xout);
  R->N = N;
  free(C.r);
  return 0;
}
