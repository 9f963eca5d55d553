// wf_fft2.h -- batched mixed-radix Stockham FFT on LDS lines for the large
// level transitions (radices 8, 4, 2, 3, 5; lengths up to 4096).
//
// Per stage of radix R with Ns = product of the radices already applied,
// butterfly b of a line (b < n/R, k = b mod Ns, h = b div Ns) reads
// v[r] = in[b + r n/R], multiplies v[r] by w_{Ns R}^{r k}, does an R-point DFT
// and writes out[h Ns R + k + r Ns].  Every stage's twiddles
// w_{Ns R}^{r k} (r = 1..R-1, k < Ns) sit in one table (built on the host in
// FP64), copied to LDS once per block; quotients by per-stage constants use
// exact multiply-high reciprocals (numerators < 2^24), so the inner loops have
// no divisions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_common.h"

#define WF_FFT2_MAX_STAGES 12

struct Fft2Plan {
  int n;
  int nstages;
  uint64_t radix_packed;  // 4 bits per stage
  const float2* tw;       // device table, twn entries (forward sign)
  int twn;
  uint32_t mag_n;                       // fmagic(n)
  uint32_t mag_ns[WF_FFT2_MAX_STAGES];  // fmagic(Ns) of each stage
};

__host__ __device__ inline int fft2_radix(const Fft2Plan& p, int s) {
  return (int)((p.radix_packed >> (4 * s)) & 15u);
}

// floor(x / d) for x < 2^24 and d >= 2 with magic = ceil(2^32 / d); d = 1
// (whose magic does not fit 32 bits) is passed as magic 0 and returns x.
__device__ inline uint32_t fdiv(uint32_t x, uint32_t magic) {
  return magic ? __umulhi(x, magic) : x;
}
__host__ __device__ inline uint32_t fmagic(uint32_t d) {
  return d <= 1 ? 0u : (uint32_t)((0x100000000ULL + d - 1) / d);
}

namespace wffft2 {

// forward R-point DFT (exp(-2 pi i q r / R)); inverse uses conjugation around it
template <int R>
__device__ inline void dft(float2* v) {
  if constexpr (R == 2) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    const float2 b0 = cadd(v[1], v[3]), b1 = csub(v[1], v[3]);
    const float2 b1r = make_float2(b1.y, -b1.x);  // -i b1
    v[0] = cadd(a0, b0);
    v[2] = csub(a0, b0);
    v[1] = cadd(a1, b1r);
    v[3] = csub(a1, b1r);
  } else if constexpr (R == 8) {
    // X[2k] = DFT4(x[n] + x[n+4]), X[2k+1] = DFT4((x[n] - x[n+4]) w8^n)
    const float h = 0.70710678118654752440f;  // 1/sqrt(2)
    float2 a[4], b[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      a[n] = cadd(v[n], v[n + 4]);
      b[n] = csub(v[n], v[n + 4]);
    }
    b[1] = make_float2(h * (b[1].x + b[1].y), h * (b[1].y - b[1].x));   // (1 - i)/sqrt2
    b[2] = make_float2(b[2].y, -b[2].x);                                 // -i
    b[3] = make_float2(h * (b[3].y - b[3].x), -h * (b[3].x + b[3].y));  // (-1 - i)/sqrt2
    dft<4>(a);
    dft<4>(b);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      v[2 * k] = a[k];
      v[2 * k + 1] = b[k];
    }
  } else if constexpr (R == 3) {
    // c = cos(2 pi/3) = -1/2, s = sin(2 pi/3) = sqrt(3)/2
    const float s3 = 0.86602540378443864676f;
    const float2 t = cadd(v[1], v[2]);
    const float2 d = csub(v[1], v[2]);
    const float2 m = make_float2(v[0].x - 0.5f * t.x, v[0].y - 0.5f * t.y);
    const float2 j = make_float2(s3 * d.y, -s3 * d.x);  // -i s d
    v[0] = cadd(v[0], t);
    v[1] = cadd(m, j);
    v[2] = csub(m, j);
  } else {  // R == 5
    // c1 = cos(2 pi/5), c2 = cos(4 pi/5), s1 = sin(2 pi/5), s2 = sin(4 pi/5)
    const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
    const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
    const float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
    const float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
    const float2 m1 = make_float2(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    const float2 m2 = make_float2(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    // -i (s1 b1 + s2 b2) and -i (s2 b1 - s1 b2)
    const float2 n1 = make_float2(s1 * b1.y + s2 * b2.y, -(s1 * b1.x + s2 * b2.x));
    const float2 n2 = make_float2(s2 * b1.y - s1 * b2.y, -(s2 * b1.x - s1 * b2.x));
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(m1, n1);
    v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2);
    v[3] = csub(m2, n2);
  }
}

template <int R>
__device__ inline void stage(const float2* __restrict__ src, float2* __restrict__ dst, int n,
                             int nlines, int lstride, int Ns, const float2* __restrict__ tw,
                             int inverse) {
  const int m = n / R;
  const uint32_t mag_m = fmagic((uint32_t)m), mag_ns = fmagic((uint32_t)Ns);
  const int total = nlines * m;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = (int)fdiv((uint32_t)t, mag_m);
    const int b = t - line * m;
    const int h = (int)fdiv((uint32_t)b, mag_ns);
    const int k = b - h * Ns;
    const float2* s = src + line * lstride;
    float2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const float2 x = s[b + r * m];
      v[r] = inverse ? make_float2(x.x, -x.y) : x;  // inverse via conjugation
    }
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[k * (R - 1) + r - 1]);
    }
    dft<R>(v);
    float2* d = dst + line * lstride + h * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[r * Ns] = inverse ? make_float2(v[r].x, -v[r].y) : v[r];
  }
}

// Batched FFT of nlines lines (length p.n, stride lstride) held in a; b is
// scratch of the same shape; tw is the plan's table in LDS.  Returns the
// buffer holding the result.  Contains barriers: call from every thread.
__device__ inline float2* lines(float2* a, float2* b, int nlines, int lstride, const Fft2Plan& p,
                                const float2* tw, int inverse) {
  int Ns = 1, off = 0;
  for (int s = 0; s < p.nstages; ++s) {
    const int R = fft2_radix(p, s);
    switch (R) {
      case 4: stage<4>(a, b, p.n, nlines, lstride, Ns, tw + off, inverse); break;
      case 8: stage<8>(a, b, p.n, nlines, lstride, Ns, tw + off, inverse); break;
      case 2: stage<2>(a, b, p.n, nlines, lstride, Ns, tw + off, inverse); break;
      case 3: stage<3>(a, b, p.n, nlines, lstride, Ns, tw + off, inverse); break;
      default: stage<5>(a, b, p.n, nlines, lstride, Ns, tw + off, inverse); break;
    }
    __syncthreads();
    off += Ns * (R - 1);
    float2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

}  // namespace wffft2

// ------------------------------------------------------------------------
// Wave-private lines (lengths up to NMAX <= 1024).  One wave transforms one
// line in its own LDS buffer; stages are in place: every lane first reads
// and transforms all of its butterflies into registers, then all lanes
// write.  LDS requests of one wave complete in issue order, so the phases
// need only a compiler barrier between them, no workgroup barrier.  The
// first forward stage reads its inputs through a functor (global loads, no
// LDS round trip); the first inverse stage forms the resampled spectrum
// from the forward spectrum on the fly; the last inverse stage hands its
// outputs (e = t + r n/R for butterfly t of the lane) to a functor in
// registers.  Inverse transforms conjugate the input and the output around
// the forward stages.
namespace wffft3 {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// butterflies per lane of a radix-R stage on lines up to NMAX (plans use
// radix 8 only on lines up to 512: one butterfly per lane)
template <int R, int NMAX>
struct Im {
  static constexpr int v = R == 8 ? 1 : (NMAX / R + 63) / 64;
};
// output registers per lane of a last stage of radix RL
template <int NMAX, int RL>
struct Slots {
  static constexpr int v = Im<RL, NMAX>::v * RL;
};

__device__ __forceinline__ float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

template <int R, int NMAX, class Src>
__device__ __forceinline__ void first_t(float2* buf, int n, Src& src, int lane) {
  constexpr int IM = Im<R, NMAX>::v;
  const int m = n / R;
  float2 v[IM][R];
  // loads under a wave-uniform guard only (lanes past m re-read butterfly
  // m - 1), so they issue back to back instead of one divergent branch each
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    if (64 * i < m) {
      const int t = min(lane + 64 * i, m - 1);
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = src(t + r * m);
    }
  }
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
      wffft2::dft<R>(v[i]);
#pragma unroll
      for (int r = 0; r < R; ++r) buf[t * R + r] = v[i][r];
    }
  }
  wave_sync();
}

template <int R, int NMAX>
__device__ __forceinline__ void mid_t(float2* buf, int n, int Ns, uint32_t mag, const float2* tw,
                                      int lane) {
  constexpr int IM = Im<R, NMAX>::v;
  const int m = n / R;
  float2 v[IM][R];
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
      const int k = t - (int)fdiv((uint32_t)t, mag) * Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[t + r * m];
#pragma unroll
      for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], tw[k * (R - 1) + r - 1]);
      wffft2::dft<R>(v[i]);
    }
  }
  wave_sync();
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
      const int h = (int)fdiv((uint32_t)t, mag), k = t - h * Ns;
      float2* d = buf + h * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) d[r * Ns] = v[i][r];
    }
  }
  wave_sync();
}

// first inverse stage: v = conj(Y[t + r m]), Y[q] = sum over f = q (mod np)
// in the band of the nc-point spectrum X (in buf) of fac[f - fmin] X[f mod nc]
template <int R, int NMAX>
__device__ __forceinline__ void rs_first_t(float2* buf, int np, int nc, uint32_t mag_np,
                                           const float2* fac, int lane) {
  constexpr int IM = Im<R, NMAX>::v;
  const int m = np / R;
  const int fmin = -(nc / 2), fmax = (nc - 1) / 2 + ((nc % 2 == 0) ? 1 : 0);
  float2 v[IM][R];
  if (np > nc) {
    // upsampling: at most one band frequency per output bin, f = q (q <=
    // fmax) or f = q - np (q - np >= fmin); np = nc folds the two halves of
    // an even length's Nyquist bin into one and takes the general path
#pragma unroll
    for (int i = 0; i < IM; ++i) {
      const int t = lane + 64 * i;
      if (t < m) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int q = t + r * m;
          const bool lo = q <= fmax, hi = q - np >= fmin;
          const int f = lo ? q : q - np;
          const int bin = f < 0 ? f + nc : f;
          float2 x = make_float2(0.f, 0.f);
          if (lo || hi) x = cmul(buf[bin], fac[f - fmin]);
          v[i][r] = conjf2(x);
        }
        wffft2::dft<R>(v[i]);
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < IM; ++i) {
      const int t = lane + 64 * i;
      if (t < m) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int q = t + r * m;
          const int f0 = q - np * (int)fdiv((uint32_t)(q - fmin), mag_np);
          float2 acc = make_float2(0.f, 0.f);
          for (int f = f0; f <= fmax; f += np) {
            const int bin = f < 0 ? f + nc : f;
            acc = cadd(acc, cmul(buf[bin], fac[f - fmin]));
          }
          v[i][r] = conjf2(acc);
        }
        wffft2::dft<R>(v[i]);
      }
    }
  }
  wave_sync();
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[t * R + r] = v[i][r];
    }
  }
  wave_sync();
}

// last forward stage (Ns R = nc, so h = 0, k = t) of a strictly upsampling
// resample (np > nc): bin b = t + r m goes straight to its padded position,
// conj(fac[f - fmin] X[b]) at q = f mod np (f = b or b - nc in the band; the
// even length's Nyquist bin to both q = nc/2 and np - nc/2), zeros between
// the band's ends; the inverse transform then starts with a plain stage
template <int R, int NMAX>
__device__ __forceinline__ void mid_rs_t(float2* buf, int nc, const float2* tw, int np,
                                         const float2* fac, int lane) {
  constexpr int IM = Im<R, NMAX>::v;
  const int m = nc / R;
  const int fmin = -(nc / 2), fmax = (nc - 1) / 2 + ((nc % 2 == 0) ? 1 : 0);
  float2 v[IM][R];
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[t + r * m];
#pragma unroll
      for (int r = 1; r < R; ++r) v[i][r] = cmul(v[i][r], tw[t * (R - 1) + r - 1]);
      wffft2::dft<R>(v[i]);
    }
  }
  wave_sync();
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int b = t + r * m;
        const int f = b <= fmax ? b : b - nc;
        buf[f >= 0 ? f : f + np] = conjf2(cmul(v[i][r], fac[f - fmin]));
        if (b == fmax && fmax == -fmin)  // even nc: the Nyquist bin's other half
          buf[np + fmin] = conjf2(cmul(v[i][r], fac[0]));
      }
    }
  }
  for (int q = fmax + 1 + lane; q < np + fmin; q += 64) buf[q] = make_float2(0.f, 0.f);
  wave_sync();
}

// first stage of a transform whose input is already in buf (Ns = 1)
template <int R, int NMAX>
__device__ __forceinline__ void plain_first_t(float2* buf, int n, int lane) {
  constexpr int IM = Im<R, NMAX>::v;
  const int m = n / R;
  float2 v[IM][R];
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[i][r] = buf[t + r * m];
      wffft2::dft<R>(v[i]);
    }
  }
  wave_sync();
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
#pragma unroll
      for (int r = 0; r < R; ++r) buf[t * R + r] = v[i][r];
    }
  }
  wave_sync();
}

// last stage (Ns R = n): outputs e = t + r m, conjugated (inverse), to sink
template <int R, int NMAX, class Sink>
__device__ __forceinline__ void last_t(const float2* buf, int n, const float2* tw, Sink& sink,
                                       int lane) {
  constexpr int IM = Im<R, NMAX>::v;
  const int m = n / R;
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
      float2 v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = buf[t + r * m];
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[t * (R - 1) + r - 1]);
      wffft2::dft<R>(v);
#pragma unroll
      for (int r = 0; r < R; ++r) sink(i * R + r, t + r * m, conjf2(v[r]));
    }
  }
  wave_sync();
}

template <int R, int NMAX, class F>
__device__ __forceinline__ void outputs_t(int n, F& f, int lane) {
  constexpr int IM = Im<R, NMAX>::v;
  const int m = n / R;
#pragma unroll
  for (int i = 0; i < IM; ++i) {
    const int t = lane + 64 * i;
    if (t < m) {
#pragma unroll
      for (int r = 0; r < R; ++r) f(i * R + r, t + r * m);
    }
  }
}

#define WF_FFT3_RADIX(R_, CALL) \
  switch (R_) {                 \
    case 4: CALL(4); break;     \
    case 8: CALL(8); break;     \
    case 2: CALL(2); break;     \
    case 3: CALL(3); break;     \
    default: CALL(5); break;    \
  }

// resample fused into the forward transform's last stage: strict
// upsampling onto a line of np, a forward plan of at least two stages
__host__ __device__ inline bool fused_resample(const Fft2Plan& pc, int np) {
  return np > pc.n && pc.nstages >= 2;
}

// Forward FFT of the line src(j), j < p.n, into buf (natural order); with
// fac (np > p.n, fused_resample) the last stage writes the conjugated,
// resampled spectrum of length np instead (inverse(..., pre = true) next).
template <int NMAX, class Src>
__device__ __forceinline__ void forward(const Fft2Plan& p, float2* buf, const float2* tw, Src& src,
                                        int lane, const float2* fac = nullptr, int np = 0) {
  const int R0 = fft2_radix(p, 0);
#define WF_C(R) first_t<R, NMAX>(buf, p.n, src, lane)
  WF_FFT3_RADIX(R0, WF_C)
#undef WF_C
  int Ns = R0, off = R0 - 1;
  const int last = fac ? p.nstages - 1 : p.nstages;
  for (int s = 1; s < last; ++s) {
    const int R = fft2_radix(p, s);
#define WF_C(R) mid_t<R, NMAX>(buf, p.n, Ns, p.mag_ns[s], tw + off, lane)
    WF_FFT3_RADIX(R, WF_C)
#undef WF_C
    off += Ns * (R - 1);
    Ns *= R;
  }
  if (fac) {
    const int R = fft2_radix(p, p.nstages - 1);
#define WF_C(R) mid_rs_t<R, NMAX>(buf, p.n, tw + off, np, fac, lane)
    WF_FFT3_RADIX(R, WF_C)
#undef WF_C
  }
}

// Spectrum of length nc in buf -> resample (fac) -> inverse FFT of length
// p.n (p.nstages >= 2, last radix RL) -> sink(slot, e, value).  pre: buf
// already holds the conjugated resampled spectrum (forward with fac).
template <int NMAX, int RL, class Sink>
__device__ __forceinline__ void inverse(const Fft2Plan& p, float2* buf, const float2* tw,
                                        const float2* fac, int nc, Sink& sink, int lane,
                                        bool pre = false) {
  const int R0 = fft2_radix(p, 0);
  if (pre) {
#define WF_C(R) plain_first_t<R, NMAX>(buf, p.n, lane)
    WF_FFT3_RADIX(R0, WF_C)
#undef WF_C
  } else {
#define WF_C(R) rs_first_t<R, NMAX>(buf, p.n, nc, p.mag_n, fac, lane)
    WF_FFT3_RADIX(R0, WF_C)
#undef WF_C
  }
  int Ns = R0, off = R0 - 1;
  for (int s = 1; s < p.nstages - 1; ++s) {
    const int R = fft2_radix(p, s);
#define WF_C(R) mid_t<R, NMAX>(buf, p.n, Ns, p.mag_ns[s], tw + off, lane)
    WF_FFT3_RADIX(R, WF_C)
#undef WF_C
    off += Ns * (R - 1);
    Ns *= R;
  }
  last_t<RL, NMAX>(buf, p.n, tw + off, sink, lane);
}

// f(slot, e) over the outputs the lane holds after inverse()
template <int NMAX, int RL, class F>
__device__ __forceinline__ void for_outputs(const Fft2Plan& p, F& f, int lane) {
  outputs_t<RL, NMAX>(p.n, f, lane);
}

}  // namespace wffft3
