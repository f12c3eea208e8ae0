// fft_lds.h -- workgroup-cooperative mixed-radix Stockham FFT in LDS.
//
// A workgroup transforms `nseq` independent complex sequences of length L that
// sit contiguously in LDS.  Stages use radix 4, 2, 3 or 5 (L = 2^a 3^b 5^c);
// each stage reads one LDS buffer and writes the other (autosort Stockham),
// so no bit reversal pass is needed.  Twiddles come from a per-length table
// in global memory: tw[t] = exp(-2 pi i t / L), conjugated for the backward
// transform.  Transforms are unnormalised, as FFTW's.
#pragma once

#include "common.h"

namespace mf {

struct FFTPlan {
  int L;
  int nstage;
  int radix[24];
};

inline bool make_fft_plan(int L, FFTPlan* P) {
  P->L = L;
  P->nstage = 0;
  int m = L;
  while (m % 4 == 0) { P->radix[P->nstage++] = 4; m /= 4; }
  while (m % 2 == 0) { P->radix[P->nstage++] = 2; m /= 2; }
  while (m % 3 == 0) { P->radix[P->nstage++] = 3; m /= 3; }
  while (m % 5 == 0) { P->radix[P->nstage++] = 5; m /= 5; }
  return m == 1 && P->nstage <= 24;
}

template <typename R, int RAD>
__device__ inline void bfly(cplx<R> (&v)[RAD], int sign) {
  using C = cplx<R>;
  if constexpr (RAD == 2) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (RAD == 4) {
    C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    C s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
    // w = exp(sign*2*pi*i/4) = sign*i ; w*d = sign*(-d.y, d.x)
    C wd = mk<R>(-(R)sign * d13.y, (R)sign * d13.x);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, wd);
    v[3] = csub(d02, wd);
  } else if constexpr (RAD == 3) {
    const R h = (R)0.86602540378443864676;  // sqrt(3)/2
    C t1 = cadd(v[1], v[2]);
    C t2 = mk<R>(v[0].x - (R)0.5 * t1.x, v[0].y - (R)0.5 * t1.y);
    C d = csub(v[1], v[2]);
    C t3 = mk<R>(-(R)sign * h * d.y, (R)sign * h * d.x);  // sign*i*h*d
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
  } else {  // RAD == 5
    const R c1 = (R)0.30901699437494742410, c2 = (R)-0.80901699437494742410;
    const R s1 = (R)0.95105651629515357212, s2 = (R)0.58778525229247312917;
    C a = v[0];
    C s14 = cadd(v[1], v[4]), d14 = csub(v[1], v[4]);
    C s23 = cadd(v[2], v[3]), d23 = csub(v[2], v[3]);
    // X1 = a + c1 s14 + c2 s23 + sign*i*(s1 d14 + s2 d23)
    // X2 = a + c2 s14 + c1 s23 + sign*i*(s2 d14 - s1 d23)
    C r1 = mk<R>(a.x + c1 * s14.x + c2 * s23.x, a.y + c1 * s14.y + c2 * s23.y);
    C r2 = mk<R>(a.x + c2 * s14.x + c1 * s23.x, a.y + c2 * s14.y + c1 * s23.y);
    C i1 = mk<R>(s1 * d14.x + s2 * d23.x, s1 * d14.y + s2 * d23.y);
    C i2 = mk<R>(s2 * d14.x - s1 * d23.x, s2 * d14.y - s1 * d23.y);
    C j1 = mk<R>(-(R)sign * i1.y, (R)sign * i1.x);
    C j2 = mk<R>(-(R)sign * i2.y, (R)sign * i2.x);
    v[0] = cadd(a, cadd(s14, s23));
    v[1] = cadd(r1, j1);
    v[4] = csub(r1, j1);
    v[2] = cadd(r2, j2);
    v[3] = csub(r2, j2);
  }
}

// one Stockham stage of radix RAD (compile time, so the butterfly stays in
// registers): src -> dst for nseq sequences of length L, Ns = product of the
// radices of the earlier stages
template <typename R, int RAD>
__device__ inline void fft_stage(const cplx<R>* src, cplx<R>* dst, int nseq, int L, int Ns,
                                 const cplx<R>* __restrict__ tw, int sign) {
  using C = cplx<R>;
  const int nbf = L / RAD;
  const int span = Ns * RAD;
  const int tscale = L / span;
  const int total = nbf * nseq;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int q = t / nbf;
    const int j = t - q * nbf;
    const C* in = src + q * L;
    C* out = dst + q * L;
    const int k = j % Ns;
    C v[RAD];
    v[0] = in[j];
#pragma unroll
    for (int r = 1; r < RAD; r++) {
      C w = tw[r * k * tscale];
      if (sign > 0) w.y = -w.y;
      v[r] = cmul(in[j + r * nbf], w);
    }
    bfly<R, RAD>(v, sign);
    const int base = (j / Ns) * span + k;
#pragma unroll
    for (int r = 0; r < RAD; r++) out[base + r * Ns] = v[r];
  }
}

// Transforms nseq sequences held in buf0 (length L each, stride L).  buf1 is
// scratch of the same size.  Returns the buffer holding the result.  Must be
// called by every thread of the workgroup; ends with a barrier.
template <typename R>
__device__ cplx<R>* lds_fft(cplx<R>* buf0, cplx<R>* buf1, int nseq, const FFTPlan& P,
                            const cplx<R>* __restrict__ tw, int sign) {
  using C = cplx<R>;
  const int L = P.L;
  C* src = buf0;
  C* dst = buf1;
  int Ns = 1;
  for (int s = 0; s < P.nstage; s++) {
    const int rad = P.radix[s];
    switch (rad) {
      case 4: fft_stage<R, 4>(src, dst, nseq, L, Ns, tw, sign); break;
      case 2: fft_stage<R, 2>(src, dst, nseq, L, Ns, tw, sign); break;
      case 3: fft_stage<R, 3>(src, dst, nseq, L, Ns, tw, sign); break;
      default: fft_stage<R, 5>(src, dst, nseq, L, Ns, tw, sign); break;
    }
    __syncthreads();
    C* tmp = src;
    src = dst;
    dst = tmp;
    Ns *= rad;
  }
  return src;
}

}  // namespace mf
