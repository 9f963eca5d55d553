// wf_fft.h -- LDS-resident mixed-radix Stockham FFT (radices 2, 3, 4, 5, 7, 11, 13)
// used by the angular interpolation of the tree aggregation.  Replaces the
// cuFFT C2C plans of treePO.cu:394-455 / interpolation.h:18-73.
//
// Stockham autosort (no bit reversal): at a stage of radix R with Ns = product
// of the radices already applied, butterfly j of a length-n line reads
// v[r] = in[j + r n/R], twiddles v[r] by w_{Ns R}^{r (j mod Ns)}, does an
// R-point DFT and writes out[(j/Ns) Ns R + (j mod Ns) + r Ns].  All twiddles
// come from one FP64-accurate table tw[m] = exp(-2 pi i m / n) per length.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_common.h"

#define WF_FFT_MAX_STAGES 16

// Radices are packed 4 bits per stage (stage s at bits 4s..4s+3) so that the
// schedule stays in scalar registers (no indexed private array).
struct FftPlanDev {
  int n;
  int nstages;
  uint64_t radix_packed;
  const float2* tw;  // n entries, exp(-2 pi i m/n)
};

__host__ __device__ inline int fft_radix(const FftPlanDev& p, int s) {
  return (int)((p.radix_packed >> (4 * s)) & 15u);
}

// Twiddle exp(sign * 2 pi i m/n) with sign = -1 (forward) or +1 (inverse).
__device__ inline float2 fft_tw(const float2* tw, int m, int inverse) {
  const float2 t = tw[m];
  return inverse ? make_float2(t.x, -t.y) : t;
}

// In-place R-point DFT of v (forward: exp(-2 pi i q r/R)).
template <int R>
__device__ inline void dft_small(float2* v, const float2* tw, int step /* n/R */, int inverse) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const float2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
    // multiply b1 by -i (forward) or +i (inverse)
    const float2 b1r = inverse ? make_float2(-b1.y, b1.x) : make_float2(b1.y, -b1.x);
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1r);
    v[3] = csub(a1, b1r);
  } else {
    float2 o[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      float2 s = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) s = cadd(s, cmul(v[r], fft_tw(tw, ((q * r) % R) * step, inverse)));
      o[q] = s;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = o[q];
  }
}

template <int R>
__device__ inline void fft_stage(const float2* __restrict__ src, float2* __restrict__ dst, int n,
                                 int nlines, int lstride, int Ns, const float2* tw, int inverse) {
  const int m = n / R;
  const int total = nlines * m;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = t / m, j = t - line * m;
    const float2* s = src + line * lstride;
    float2* d = dst + line * lstride;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[j + r * m];
    const int k = j % Ns;
    if (Ns > 1) {
      const int tstep = n / (Ns * R);  // w_{Ns R}^{r k} = tw[r k n/(Ns R)]
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], fft_tw(tw, r * k * tstep, inverse));
    }
    dft_small<R>(v, tw, m, inverse);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[base + r * Ns] = v[r];
  }
}

// Batched FFT of nlines lines of length p.n stored at a + line*lstride.
// Ping-pongs between a and b; returns the buffer holding the result.
// Must be called by every thread of the block (contains barriers).
__device__ inline float2* fft_lines(float2* a, float2* b, int nlines, int lstride,
                                    const FftPlanDev& p, int inverse) {
  int Ns = 1;
  for (int s = 0; s < p.nstages; ++s) {
    const int R = fft_radix(p, s);
    switch (R) {
      case 2: fft_stage<2>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 3: fft_stage<3>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 4: fft_stage<4>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 5: fft_stage<5>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 7: fft_stage<7>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 11: fft_stage<11>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      default: fft_stage<13>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
    }
    __syncthreads();
    float2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// Spectral resample of one set of lines: input spectra X (length nc, as
// produced by the forward FFT), output Y (length np) with
//   Y[q] = (1/nc) sum_{f = q mod np, f in band(nc)} X[f mod nc] exp(i f shift)
// band(nc) = signed frequencies of the trigonometric interpolant, the Nyquist
// bin (nc even) split in half between +nc/2 and -nc/2 (interpolation.h:2-9,
// :24-37).  np >= nc is the reference's zero-pad; np < nc folds (exact
// evaluation of the interpolant on the coarser grid).
__device__ inline void spectral_resample(const float2* __restrict__ X, float2* __restrict__ Y,
                                         int nlines, int xstride, int ystride, int nc, int np,
                                         float shift, float scale) {
  const int fmin = -(nc / 2), fmax = (nc - 1) / 2 + ((nc % 2 == 0) ? 1 : 0);
  const int total = nlines * np;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = t / np, q = t - line * np;
    const float2* x = X + line * xstride;
    // smallest f >= fmin with f = q (mod np)
    int f = q - ((q - fmin) / np) * np;
    if (f < fmin) f += np;
    float2 acc = make_float2(0.f, 0.f);
    for (; f <= fmax; f += np) {
      const int bin = f < 0 ? f + nc : f;
      float2 v = x[bin];
      if ((nc % 2 == 0) && (f == nc / 2 || f == -(nc / 2))) v = cscale(v, 0.5f);
      if (shift != 0.f) {
        float s, c;
        sincosf(shift * (float)f, &s, &c);
        v = cmul(v, make_float2(c, s));
      }
      acc = cadd(acc, v);
    }
    Y[line * ystride + q] = cscale(acc, scale);
  }
}

// Hermitian adjoint of spectral_resample (without its scale): for every
// child bin m, gather the parent bins its band frequencies map to,
//   X[m] = sum_{f in band(nc), f = m mod nc} w_f exp(-i f shift) Y[f mod np],
// w_f = 1/2 for the two Nyquist halves (nc even), else 1.
__device__ inline void spectral_resample_adj(const float2* __restrict__ Y, float2* __restrict__ X,
                                             int nlines, int ystride, int xstride, int nc, int np,
                                             float shift, float scale) {
  const int total = nlines * nc;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = t / nc, m = t - line * nc;
    const float2* y = Y + line * ystride;
    float2 acc = make_float2(0.f, 0.f);
    const bool nyq = (nc % 2 == 0) && (m == nc / 2);
    const int nf = nyq ? 2 : 1;
    for (int h = 0; h < nf; ++h) {
      int f = (m <= (nc - 1) / 2) ? m : m - nc;
      if (nyq) f = h ? -(nc / 2) : nc / 2;
      int q = f % np;
      if (q < 0) q += np;
      float2 v = y[q];
      if (nyq) v = cscale(v, 0.5f);
      if (shift != 0.f) {
        float s, c;
        sincosf(-shift * (float)f, &s, &c);
        v = cmul(v, make_float2(c, s));
      }
      acc = cadd(acc, v);
    }
    X[line * xstride + m] = cscale(acc, scale);
  }
}
