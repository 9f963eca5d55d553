// wf_agg_fft.hip -- level transitions with large angular grids through the
// LDS FFT of wf_fft2.h.
//
// For long lines the dense interpolation GEMM has to re-read either the data
// or its (rows x columns) matrix many times; the FFT route reads each child
// sample once and writes each parent sample once:
//
//   phi_fft:   per block RB theta rows of one (child, field): FFT(npc) ->
//              resample to npp (pad / fold, Nyquist halved, 1/npc) ->
//              inverse FFT(npp) -> GT[cf][q][theta] (theta contiguous,
//              stride fft_g_stride(ntc)).
//   theta_fft: per block W column pairs of one (parent, field); for each
//              child the great-circle sequences (column k ascending, column
//              k + npp/2 reversed) -> FFT(2 ntc) -> resample with the origin
//              shift -> inverse FFT(2 ntp) -> times the translation table of
//              the child's octant -> parent accumulators in registers (FP64
//              at the root).  Same extended-row mapping as the MFMA path.
#include "wf_common.h"
#include "wf_fft2.h"
#include "wf_interp.h"

namespace {

constexpr int kLineBuf = 4096;  // complex elements per LDS ping-pong buffer

template <bool ROOT>
struct FftAcc {
  typedef float2 type;
};
template <>
struct FftAcc<true> {
  typedef double2 type;
};

// Spectral resample of nlines lines: X (length nc spectra, stride xs) -> Y
// (length np, stride ys), Y[q] = sum_{f = q mod np, f in band(nc)} fac[f - fmin] X[f mod nc]
// with fac[] = w_f exp(i f shift) / nc precomputed in LDS.
__device__ inline void resample(const float2* __restrict__ X, float2* __restrict__ Y, int nlines,
                                int xs, int ys, int nc, int np, const float2* __restrict__ fac) {
  const int fmin = -(nc / 2), fmax = (nc - 1) / 2 + ((nc % 2 == 0) ? 1 : 0);
  const uint32_t mag = fmagic((uint32_t)np);
  const int total = nlines * np;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = (int)fdiv((uint32_t)t, mag);
    const int q = t - line * np;
    const float2* x = X + line * xs;
    // smallest f >= fmin with f = q (mod np): f = q - np * floor((q - fmin) / np)
    const int f0 = q - np * (int)fdiv((uint32_t)(q - fmin), mag);
    float2 acc = make_float2(0.f, 0.f);
    for (int f = f0; f <= fmax; f += np) {
      const int bin = f < 0 ? f + nc : f;
      acc = cadd(acc, cmul(x[bin], fac[f - fmin]));
    }
    Y[line * ys + q] = acc;
  }
}

__device__ inline void load_table(float2* dst, const float2* src, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = src[t];
}

template <int NT>
__global__ __launch_bounds__(NT) void phi_fft_kernel(const float2* __restrict__ X,
                                                      float2* __restrict__ GT, int ntc, int npc,
                                                      int npp, int RB, Fft2Plan pc, Fft2Plan pp,
                                                      const float2* __restrict__ fac_g) {
  extern __shared__ float2 sm[];
  const int lmax = npc > npp ? npc : npp;
  float2* a = sm;
  float2* b = a + RB * lmax;
  float2* twc = b + RB * lmax;
  float2* twp = twc + pc.twn;
  float2* fac = twp + pp.twn;
  load_table(twc, pc.tw, pc.twn);
  load_table(twp, pp.tw, pp.twn);
  load_table(fac, fac_g, npc + 1);
  const int rblocks = (ntc + RB - 1) / RB;
  const int64_t cf = blockIdx.x / rblocks;
  const int r0 = (int)(blockIdx.x - cf * rblocks) * RB;
  const int nr = min(RB, ntc - r0);
  const float2* src = X + (cf * ntc + r0) * (int64_t)npc;
  for (int t = threadIdx.x; t < nr * npc; t += blockDim.x) {
    const int r = t / npc, k = t - r * npc;
    a[r * lmax + k] = src[t];
  }
  __syncthreads();
  float2* Xs = wffft2::lines(a, b, nr, lmax, pc, twc, 0);
  float2* Ys = (Xs == a) ? b : a;
  resample(Xs, Ys, nr, lmax, lmax, npc, npp, fac);
  __syncthreads();
  float2* Z = wffft2::lines(Ys, Xs, nr, lmax, pp, twp, 1);
  // GT[cf][q][r0 + r]: consecutive r contiguous
  const int ldg = wf::fft_g_stride(ntc);
  float2* dst = GT + cf * (int64_t)npp * ldg + r0;
  for (int t = threadIdx.x; t < nr * npp; t += blockDim.x) {
    const int q = t / nr, r = t - q * nr;
    dst[(int64_t)q * ldg + r] = Z[r * lmax + q];
  }
}

template <int MAXS, bool ROOT, int NT>
__global__ __launch_bounds__(NT) void theta_fft_kernel(
    const float2* __restrict__ GT, int F, int ntc, int npp, int ntp, int W,
    const int32_t* __restrict__ cb, const int32_t* __restrict__ ce,
    const uint32_t* __restrict__ child_key, const float2* __restrict__ Text, int text_per_child,
    int mirror, Fft2Plan pc, Fft2Plan pp, const float2* __restrict__ fac_g,
    float2* __restrict__ out, double2* __restrict__ out_root, int flip) {
  typedef typename FftAcc<ROOT>::type acc_t;
  // transverse components: the second half of every great circle is negated
  // on input and on output
  const float sgn2 = flip ? -1.f : 1.f;
  extern __shared__ float2 sm[];
  const int n2c = 2 * ntc, n2p = 2 * ntp, half = npp / 2;
  const int lmax = n2c > n2p ? n2c : n2p;
  float2* a = sm;
  float2* b = a + W * lmax;
  float2* twc = b + W * lmax;
  float2* twp = twc + pc.twn;
  float2* fac = twp + pp.twn;
  load_table(twc, pc.tw, pc.twn);
  load_table(twp, pp.tw, pp.twn);
  load_table(fac, fac_g, n2c + 1);
  const int nwb = (half + W - 1) / W;
  const int f = (int)(blockIdx.x % F);
  const int64_t rest = blockIdx.x / F;
  const int wb = (int)(rest % nwb);
  const int p = (int)(rest / nwb);
  const int k0 = wb * W;
  const int nw = min(W, half - k0);
  const int i2_0 = mirror % n2p;
  const bool dual = (i2_0 < ntp);
  // slot s = threadIdx.x + NT m -> (row e = s / nw, column w = s % nw): lanes
  // with consecutive w read the translation table and write the parent grid
  // at consecutive addresses
  const uint32_t mag_nw = fmagic((uint32_t)nw);
  // acc2: the second parent sample fed by great-circle row e = 0 (pole
  // grids).  Row 0 is slots s < nw, and nw <= W <= NT (launcher), so only
  // slot m = 0 can hold it: one extra accumulator, not MAXS.
  acc_t acc[MAXS], acc2;
#pragma unroll
  for (int m = 0; m < MAXS; ++m) {
    acc[m].x = 0;
    acc[m].y = 0;
  }
  acc2.x = 0;
  acc2.y = 0;
  const int c0 = cb ? cb[p] : p, c1 = cb ? ce[p] : p + 1;
  for (int c = c0; c < c1; ++c) {
    const int ldg = wf::fft_g_stride(ntc);
    const float2* S = GT + ((int64_t)c * F + f) * (int64_t)npp * ldg;
    __syncthreads();  // previous child done with a/b
    for (int t = threadIdx.x; t < nw * n2c; t += blockDim.x) {
      const int w = t / n2c, jp = t - w * n2c;
      const int k = k0 + w;
      if (jp < ntc) {
        a[w * lmax + jp] = S[(int64_t)k * ldg + jp];
      } else {
        const float2 v = S[(int64_t)(k + half) * ldg + (n2c - 1 - jp)];
        a[w * lmax + jp] = make_float2(sgn2 * v.x, sgn2 * v.y);
      }
    }
    __syncthreads();
    float2* Xs = wffft2::lines(a, b, nw, lmax, pc, twc, 0);
    float2* Ys = (Xs == a) ? b : a;
    resample(Xs, Ys, nw, lmax, lmax, n2c, n2p, fac);
    __syncthreads();
    float2* Z = wffft2::lines(Ys, Xs, nw, lmax, pp, twp, 1);
    const int oct = text_per_child ? c : (child_key ? (int)(child_key[c] & 7u) : 0);
    const float2* To = Text ? Text + (int64_t)oct * (n2p + 1) * half : nullptr;
#pragma unroll
    for (int m = 0; m < MAXS; ++m) {
      const int s = threadIdx.x + NT * m;
      if (s < nw * n2p) {
        const int e = (int)fdiv((uint32_t)s, mag_nw), w = s - e * nw;
        const float2 u = Z[w * lmax + e];
        const float2 tu = To ? cmul(To[(int64_t)e * half + k0 + w], u) : u;
        acc[m].x += tu.x;
        acc[m].y += tu.y;
        if (m == 0 && dual && e == 0) {
          const float2 tu2 = To ? cmul(To[(int64_t)n2p * half + k0 + w], u) : u;
          acc2.x += tu2.x;
          acc2.y += tu2.y;
        }
      }
    }
  }
  const int64_t obase = ((int64_t)p * F + f) * (int64_t)ntp * npp;
#pragma unroll
  for (int m = 0; m < MAXS; ++m) {
    const int s = threadIdx.x + NT * m;
    if (s >= nw * n2p) continue;
    const int e = (int)fdiv((uint32_t)s, mag_nw), w = s - e * nw;
    const int col = k0 + w;
    int row = e, cc = col;
    float sg = 1.f;
    if (e >= ntp) {
      const int i2 = mirror - e;  // in [0, ntp]
      row = i2 < ntp ? i2 : -1;
      cc = col + half;
      sg = sgn2;
    }
    if (row >= 0) {
      const int64_t o = obase + (int64_t)row * npp + cc;
      if (ROOT)
        out_root[o] = make_double2(sg * (double)acc[m].x, sg * (double)acc[m].y);
      else
        out[o] = make_float2(sg * (float)acc[m].x, sg * (float)acc[m].y);
    }
    if (m == 0 && dual && e == 0) {
      const int64_t o2 = obase + (int64_t)i2_0 * npp + col + half;
      if (ROOT)
        out_root[o2] = make_double2(sgn2 * (double)acc2.x, sgn2 * (double)acc2.y);
      else
        out[o2] = make_float2(sgn2 * (float)acc2.x, sgn2 * (float)acc2.y);
    }
  }
}

// ---------------------------------------------------------------- wave lines --
// phi: block = nw (8 or 16) waves on nw consecutive theta rows of one
// (child, field) plane; each wave transforms its row (FFT(npc), resample, inverse FFT(npp))
// and leaves the npp outputs in its LDS line; the block then writes
// GT[cf][q][r0 .. r0 + nw) with the rows contiguous.
template <int NMAX, int RL, bool PRE>
__global__ __launch_bounds__(1024) void phi_fft3_kernel(const float2* __restrict__ X,
                                                       float2* __restrict__ GT, int ntc, int npc,
                                                       int npp, Fft2Plan pc, Fft2Plan pp,
                                                       const float2* __restrict__ fac_g) {
  extern __shared__ float2 sm[];
  const int nw = (int)(blockDim.x >> 6), wave = (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // odd line stride: the transposed read-out (one sample of every line per
  // 16 lanes) then spreads over the LDS banks
  const int lmax = (npc > npp ? npc : npp) | 1;
  float2* twc = sm;
  float2* twp = twc + pc.twn;
  float2* fac = twp + pp.twn;
  float2* lines = fac + (npc + 1);
  load_table(twc, pc.tw, pc.twn);
  load_table(twp, pp.tw, pp.twn);
  load_table(fac, fac_g, npc + 1);
  __syncthreads();
  const int rblocks = (ntc + nw - 1) / nw;
  const int64_t cf = blockIdx.x / rblocks;
  const int r0 = (int)(blockIdx.x - cf * rblocks) * nw;
  const int nr = min(nw, ntc - r0);
  float2* buf = lines + wave * lmax;
  if (wave < nr) {
    const float2* row = X + (cf * ntc + r0 + wave) * (int64_t)npc;
    auto src = [&](int j) { return row[j]; };
    wffft3::forward<NMAX>(pc, buf, twc, src, lane, PRE ? fac : nullptr, npp);
    float2 o[wffft3::Slots<NMAX, RL>::v];
    auto keep = [&](int s, int, float2 z) { o[s] = z; };
    wffft3::inverse<NMAX, RL>(pp, buf, twp, fac, npc, keep, lane, PRE);
    auto put = [&](int s, int e) { buf[e] = o[s]; };
    wffft3::for_outputs<NMAX, RL>(pp, put, lane);
  }
  __syncthreads();
  // nw is a power of two (launcher): thread t -> (q = t / nw, row w = t % nw)
  const int ldg = wf::fft_g_stride(ntc);
  float2* dst = GT + cf * (int64_t)npp * ldg + r0;
  const int sh = __builtin_ctz((unsigned)nw);
  for (int t = threadIdx.x; t < (npp << sh); t += blockDim.x) {
    const int q = t >> sh, w = t & (nw - 1);
    if (w < nr) dst[(int64_t)q * ldg + w] = lines[w * lmax + q];
  }
}

// theta: block = kNw waves on kNw consecutive great circles (column pairs k,
// k + npp/2) of one (parent, field); each wave runs every child's line
// (FFT(2 ntc), resample with the origin shift, inverse FFT(2 ntp)), multiplies
// its outputs by the child's translation factors (k-major table: one
// contiguous row per great circle) and accumulates them in registers (FP64 at
// the root).  Block b -> (u = parent * nkg + k group, field): blocks go
// round-robin over the 8 XCDs, and XCD x takes a contiguous range of u with
// the F fields of one u consecutive, so a root child's table rows are fetched
// into that XCD's L2 once for all fields and neighbouring k groups, which
// write the same parent rows, share one L2.
constexpr int kNw = 4;
template <int NMAX, int RL, bool ROOT, bool PRE>
__global__ __launch_bounds__(64 * kNw, 3) void theta_fft3_kernel(
    const float2* __restrict__ GT, int F, int ntc, int npp, int ntp, int64_t U, int nkg,
    const int32_t* __restrict__ cb, const int32_t* __restrict__ ce,
    const uint32_t* __restrict__ child_key, const float2* __restrict__ Text, int text_per_child,
    int mirror, Fft2Plan pc, Fft2Plan pp, const float2* __restrict__ fac_g,
    float2* __restrict__ out, double2* __restrict__ out_root, int flip) {
  typedef typename FftAcc<ROOT>::type acc_t;
  extern __shared__ float2 sm[];
  const int wave = (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int n2c = 2 * ntc, n2p = 2 * ntp, half = npp / 2;
  const int lmax = n2c > n2p ? n2c : n2p;
  float2* twc = sm;
  float2* twp = twc + pc.twn;
  float2* fac = twp + pp.twn;
  float2* lines = fac + (n2c + 1);
  load_table(twc, pc.tw, pc.twn);
  load_table(twp, pp.tw, pp.twn);
  load_table(fac, fac_g, n2c + 1);
  __syncthreads();
  const int xcd = (int)(blockIdx.x & 7u);
  const int64_t j = blockIdx.x >> 3;
  const int f = (int)(j % F);
  const int64_t u = xcd * ((U + 7) / 8) + j / F;  // XCD x: k groups [x per, (x + 1) per)
  if (u >= U) return;
  const int p = (int)(u / nkg);
  const int k = (int)(u - (int64_t)p * nkg) * kNw + wave;
  if (k >= half) return;
  // transverse components: the second half of every great circle negated
  const float sgn2 = flip ? -1.f : 1.f;
  const int i2_0 = mirror % n2p;
  const bool dual = (i2_0 < ntp);
  constexpr int NS = wffft3::Slots<NMAX, RL>::v;
  acc_t acc[NS], acc2;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    acc[s].x = 0;
    acc[s].y = 0;
  }
  acc2.x = 0;
  acc2.y = 0;
  float2* buf = lines + wave * lmax;
  const int ldg = wf::fft_g_stride(ntc);
  const int c0 = cb ? cb[p] : p, c1 = cb ? ce[p] : p + 1;
  for (int c = c0; c < c1; ++c) {
    // the lane index through an empty asm: the stages' lane-derived addresses
    // are recomputed per child instead of all being hoisted out of the loop
    int ln = lane;
    __asm__ volatile("" : "+v"(ln));
    const float2* S = GT + ((int64_t)c * F + f) * (int64_t)npp * ldg;
    const float2* S0 = S + (int64_t)k * ldg;
    const float2* S1 = S + (int64_t)(k + half) * ldg + (n2c - 1);
    // branch-free gather (address and sign selected per sample), so the
    // stage's loads issue back to back and are waited for once
    auto src = [&](int jp) -> float2 {
      const bool first = jp < ntc;
      const float2 v = *(first ? S0 + jp : S1 - jp);
      const float sg = first ? 1.f : sgn2;
      return make_float2(sg * v.x, sg * v.y);
    };
    wffft3::forward<NMAX>(pc, buf, twc, src, ln, PRE ? fac : nullptr, n2p);
    const int oct = text_per_child ? c : (child_key ? (int)(child_key[c] & 7u) : 0);
    const float2* To = Text ? Text + ((int64_t)oct * half + k) * (n2p + 1) : nullptr;
    const float2 tdual = (To && dual) ? To[n2p] : make_float2(1.f, 0.f);  // pole row
    auto sink = [&](int s, int e, float2 z) {
      const float2 tu = To ? cmul(To[e], z) : z;
      acc[s].x += tu.x;
      acc[s].y += tu.y;
      if (dual && e == 0) {
        const float2 t2 = cmul(tdual, z);
        acc2.x += t2.x;
        acc2.y += t2.y;
      }
    };
    wffft3::inverse<NMAX, RL>(pp, buf, twp, fac, n2c, sink, ln, PRE);
  }
  const int64_t obase = ((int64_t)p * F + f) * (int64_t)ntp * npp;
  auto put = [&](int s, int e) {
    int row = e, cc = k;
    float sg = 1.f;
    if (e >= ntp) {
      const int i2 = mirror - e;  // in [0, ntp]
      row = i2 < ntp ? i2 : -1;
      cc = k + half;
      sg = sgn2;
    }
    if (row >= 0) {
      const int64_t o = obase + (int64_t)row * npp + cc;
      if constexpr (ROOT)
        out_root[o] = make_double2(sg * (double)acc[s].x, sg * (double)acc[s].y);
      else
        out[o] = make_float2(sg * (float)acc[s].x, sg * (float)acc[s].y);
    }
    if (dual && e == 0) {
      const int64_t o2 = obase + (int64_t)i2_0 * npp + k + half;
      if constexpr (ROOT)
        out_root[o2] = make_double2(sgn2 * (double)acc2.x, sgn2 * (double)acc2.y);
      else
        out[o2] = make_float2(sgn2 * (float)acc2.x, sgn2 * (float)acc2.y);
    }
  };
  wffft3::for_outputs<NMAX, RL>(pp, put, lane);
}

template <int NMAX, int RL, bool ROOT, bool PRE>
hipError_t launch_theta_fft3_p(const wf::ThetaFftArgs& a, hipStream_t st) {
  const int n2c = 2 * a.ntc, n2p = 2 * a.ntp, half = a.npp / 2;
  const int lmax = n2c > n2p ? n2c : n2p;
  const size_t sh = sizeof(float2) * ((size_t)a.pc.twn + a.pp.twn + n2c + 1 + (size_t)kNw * lmax);
  if (sh > 160 * 1024) return hipErrorInvalidValue;
  const int nkg = (half + kNw - 1) / kNw;
  const int64_t U = (int64_t)a.nparents * nkg;
  const int64_t nblk = ((U + 7) / 8) * 8 * a.F;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  hipError_t e = hipFuncSetAttribute((const void*)theta_fft3_kernel<NMAX, RL, ROOT, PRE>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((theta_fft3_kernel<NMAX, RL, ROOT, PRE>), dim3((unsigned)nblk),
                     dim3(64 * kNw), sh, st, a.GT, a.F, a.ntc, a.npp, a.ntp, U, nkg, a.cb, a.ce,
                     a.child_key, a.Text, a.text_per_child, a.mirror, a.pc, a.pp, a.fac, a.out,
                     a.out_root, a.flip);
  return hipGetLastError();
}

template <int NMAX, int RL, bool ROOT>
hipError_t launch_theta_fft3_t(const wf::ThetaFftArgs& a, hipStream_t st) {
  if (wffft3::fused_resample(a.pc, 2 * a.ntp))
    return launch_theta_fft3_p<NMAX, RL, ROOT, true>(a, st);
  return launch_theta_fft3_p<NMAX, RL, ROOT, false>(a, st);
}

template <bool ROOT>
hipError_t launch_theta_fft3(const wf::ThetaFftArgs& a, hipStream_t st) {
  const int lmax = a.ntc > a.ntp ? 2 * a.ntc : 2 * a.ntp;
  const int rl = fft2_radix(a.pp, a.pp.nstages - 1);
  if (rl == 4) {
    if (lmax <= 512) return launch_theta_fft3_t<512, 4, ROOT>(a, st);
    if (lmax <= 768) return launch_theta_fft3_t<768, 4, ROOT>(a, st);
    return launch_theta_fft3_t<1024, 4, ROOT>(a, st);
  }
  if (rl == 8) {
    if (lmax <= 512) return launch_theta_fft3_t<512, 8, ROOT>(a, st);
    if (lmax <= 768) return launch_theta_fft3_t<768, 8, ROOT>(a, st);
    return launch_theta_fft3_t<1024, 8, ROOT>(a, st);
  }
  if (rl != 2) return hipErrorInvalidValue;  // even lengths end on radix 8, 4 or 2
  if (lmax <= 512) return launch_theta_fft3_t<512, 2, ROOT>(a, st);
  if (lmax <= 768) return launch_theta_fft3_t<768, 2, ROOT>(a, st);
  return launch_theta_fft3_t<1024, 2, ROOT>(a, st);
}

template <bool ROOT, int NT>
hipError_t launch_theta_fft_t(const wf::ThetaFftArgs& a, int W, int slots, size_t sh,
                              hipStream_t st) {
  const int half = a.npp / 2;
  const int64_t nblk = (int64_t)a.nparents * a.F * ((half + W - 1) / W);
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
#define WF_THETA_FFT(MS)                                                                          \
  do {                                                                                            \
    hipError_t e = hipFuncSetAttribute((const void*)theta_fft_kernel<MS, ROOT, NT>,               \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);      \
    if (e != hipSuccess) return e;                                                                \
    hipLaunchKernelGGL((theta_fft_kernel<MS, ROOT, NT>), dim3((unsigned)nblk), dim3(NT), sh, st,   \
                       a.GT, a.F, a.ntc, a.npp, a.ntp, W, a.cb, a.ce, a.child_key, a.Text,         \
                       a.text_per_child, a.mirror, a.pc, a.pp, a.fac, a.out, a.out_root,          \
                       a.flip);                                                                   \
    return hipGetLastError();                                                                     \
  } while (0)
  if (slots <= 4) WF_THETA_FFT(4);
  if (slots <= 8) WF_THETA_FFT(8);
  WF_THETA_FFT(16);
#undef WF_THETA_FFT
}

// threads per block of the FFT kernels and the LDS they may use for lines.
// Half the CU's 160 KB lets two 512-thread blocks share a CU, so one block's
// FFT barriers overlap the other's work (fiber 663k root transition: -8 %
// theta time against a 150 KB budget with one block per CU; 48 KB is slower
// again, the lines per block get too few).
int fft_threads() { return 512; }
constexpr size_t kFftLds = 75 * 1024;
constexpr size_t kFftLdsMax = 150 * 1024;  // when 75 KB cannot hold the tables and one line pair

// Lines of length lmax (two ping-pong buffers each) that fit beside the
// tables: in kFftLds, or in kFftLdsMax when not even one fits there.
int fft_lines(size_t tables, int lmax) {
  const size_t per = 2 * sizeof(float2) * (size_t)lmax;
  if (tables + per <= kFftLds) return (int)((kFftLds - tables) / per);
  if (tables + per <= kFftLdsMax) return (int)((kFftLdsMax - tables) / per);
  return 1;  // the launcher's 160 KB check rejects it
}

}  // namespace

namespace wf {

hipError_t launch_phi_fft(const float2* X, float2* GT, int64_t lines_cf, int ntc, int npc, int npp,
                          const Fft2Plan& pc, const Fft2Plan& pp, const float2* fac,
                          hipStream_t st) {
  if (lines_cf <= 0) return hipSuccess;
  const int lmax = npc > npp ? npc : npp;
  if (pc.n != npc || pp.n != npp) return hipErrorInvalidValue;
  if (phi_fft_wave(npc, npp)) {
    if (pp.nstages < 2) return hipErrorInvalidValue;
    // rows per block: 16 (128 B runs of the transposed writes, full cache
    // lines) where the LDS fits two blocks per CU, else 8
    const size_t tab = (size_t)pc.twn + pp.twn + npc + 1;
    const size_t lstr = (size_t)(lmax | 1);  // the kernel's line stride
    const int NW = (ntc >= 16 && sizeof(float2) * (tab + 16 * lstr) <= 80 * 1024) ? 16 : 8;
    const size_t sh = sizeof(float2) * (tab + (size_t)NW * lstr);
    if (sh > 160 * 1024) return hipErrorInvalidValue;
    const int64_t nblk = lines_cf * ((ntc + NW - 1) / NW);
    if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
    const int rl = fft2_radix(pp, pp.nstages - 1);
    if (rl != 2 && rl != 4 && rl != 8) return hipErrorInvalidValue;  // even lengths end on 8, 4, 2
    const bool pre = wffft3::fused_resample(pc, npp);
#define WF_PHI3P(N, RL, P)                                                                        \
  do {                                                                                            \
    hipError_t e = hipFuncSetAttribute((const void*)phi_fft3_kernel<N, RL, P>,                    \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);      \
    if (e != hipSuccess) return e;                                                                \
    hipLaunchKernelGGL((phi_fft3_kernel<N, RL, P>), dim3((unsigned)nblk), dim3(64 * NW), sh, st,  \
                       X, GT, ntc, npc, npp, pc, pp, fac);                                        \
    return hipGetLastError();                                                                     \
  } while (0)
#define WF_PHI3(N, RL)         \
  do {                         \
    if (pre)                   \
      WF_PHI3P(N, RL, true);   \
    else                       \
      WF_PHI3P(N, RL, false);  \
  } while (0)
    if (rl == 4) {
      if (lmax <= 512) WF_PHI3(512, 4);
      if (lmax <= 768) WF_PHI3(768, 4);
      WF_PHI3(1024, 4);
    }
    if (rl == 8) {
      if (lmax <= 512) WF_PHI3(512, 8);
      if (lmax <= 768) WF_PHI3(768, 8);
      WF_PHI3(1024, 8);
    }
    if (lmax <= 512) WF_PHI3(512, 2);
    if (lmax <= 768) WF_PHI3(768, 2);
    WF_PHI3(1024, 2);
#undef WF_PHI3
#undef WF_PHI3P
  }
  if (lmax > kLineBuf || pc.n != npc || pp.n != npp) return hipErrorInvalidValue;
  const int NT = fft_threads();
  const size_t tables = sizeof(float2) * ((size_t)pc.twn + pp.twn + npc + 1);
  int RB = NT == 256 ? kLineBuf / lmax : fft_lines(tables, lmax);
  RB = RB < 1 ? 1 : (RB > ntc ? ntc : RB);
  const size_t sh = sizeof(float2) * (size_t)2 * RB * lmax + tables;
  if (sh > 160 * 1024) return hipErrorInvalidValue;
  const int64_t nblk = lines_cf * ((ntc + RB - 1) / RB);
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
#define WF_PHI_FFT(T)                                                                             \
  do {                                                                                            \
    hipError_t e = hipFuncSetAttribute((const void*)phi_fft_kernel<T>,                            \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);      \
    if (e != hipSuccess) return e;                                                                \
    hipLaunchKernelGGL(phi_fft_kernel<T>, dim3((unsigned)nblk), dim3(T), sh, st, X, GT, ntc, npc,  \
                       npp, RB, pc, pp, fac);                                                     \
    return hipGetLastError();                                                                     \
  } while (0)
  if (NT == 1024) WF_PHI_FFT(1024);
  if (NT == 512) WF_PHI_FFT(512);
  WF_PHI_FFT(256);
#undef WF_PHI_FFT
}

hipError_t launch_theta_fft(const ThetaFftArgs& a, hipStream_t st) {
  if (a.nparents <= 0) return hipSuccess;
  const int n2c = 2 * a.ntc, n2p = 2 * a.ntp, half = a.npp / 2;
  if (a.npp < 2 || (a.npp & 1) || a.pc.n != n2c || a.pp.n != n2p) return hipErrorInvalidValue;
  if (a.mirror != n2p - 1 && a.mirror != n2p) return hipErrorInvalidValue;
  if ((a.out_root != nullptr) == (a.out != nullptr)) return hipErrorInvalidValue;
  const int lmax = n2c > n2p ? n2c : n2p;
  if (theta_fft_wave(a.ntc, a.ntp)) {
    if (a.Text && !a.text_kmajor) return hipErrorInvalidValue;
    if (a.pp.nstages < 2) return hipErrorInvalidValue;
    if (a.out_root) return launch_theta_fft3<true>(a, st);
    return launch_theta_fft3<false>(a, st);
  }
  if (a.Text && a.text_kmajor) return hipErrorInvalidValue;
  const int NT = fft_threads();
  if (lmax > kLineBuf || n2p > NT * 16) return hipErrorInvalidValue;
  const size_t tables = sizeof(float2) * ((size_t)a.pc.twn + a.pp.twn + n2c + 1);
  int W = NT == 256 ? kLineBuf / lmax : fft_lines(tables, lmax);
  W = W < 1 ? 1 : (W > half ? half : W);
  if (W > NT) W = NT;  // the kernel keeps row 0's second sample for slot 0 only
  const int max_slots = NT == 1024 ? 8 : 16;  // accumulator registers per thread
  while (W > 1 && W * n2p > NT * max_slots) --W;
  const int slots = (W * n2p + NT - 1) / NT;
  const size_t sh = sizeof(float2) * (size_t)2 * W * lmax + tables;
  if (sh > 160 * 1024) return hipErrorInvalidValue;
  if (NT == 1024) {
    if (a.out_root) return launch_theta_fft_t<true, 1024>(a, W, slots, sh, st);
    return launch_theta_fft_t<false, 1024>(a, W, slots, sh, st);
  }
  if (NT == 512) {
    if (a.out_root) return launch_theta_fft_t<true, 512>(a, W, slots, sh, st);
    return launch_theta_fft_t<false, 512>(a, W, slots, sh, st);
  }
  if (a.out_root) return launch_theta_fft_t<true, 256>(a, W, slots, sh, st);
  return launch_theta_fft_t<false, 256>(a, W, slots, sh, st);
}

}  // namespace wf
