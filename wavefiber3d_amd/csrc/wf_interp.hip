// wf_interp.hip -- upward aggregation of the PO tree.
//
// Reference per (parent, active child, component): copyandinterp
// (interpolation.h:82-92: fill, copy, 4 cuFFT C2C, 2 (Nt_c + Np_p) row
// copies, Nyquist fix-ups, 2 scalings), translation_compute (fields.h:92-103,
// host FP32 table treePO.cu:511-554) and sumchildren (fields.h:153-170), each
// a separate launch, plus a linear host search for the child
// (treePO.cu:725-732).  Here one level transition is one fused kernel when
// its grids fit in LDS (agg_fused: one block per (parent, field), every
// child pattern read once, the parent written once), otherwise two kernels:
//
//  phi_pass:  every child row (theta_j, all phi) -> FFT(npc) -> spectral
//             pad/fold to npp -> IFFT(npp), in LDS.
//  theta_pass: one block per (parent, field, tile of W phi columns); for each
//             active child the columns phi_k and phi_k + pi form one great
//             circle (the 2 pi extension theta -> pi - theta, phi -> phi + pi;
//             exact on the half-sample child grid); FFT(2 ntc) -> shift to the
//             parent theta origin, pad/fold to 2 ntp -> IFFT(2 ntp) ->
//             multiply by the translation exp(i k s.(c_child - c_parent)) and
//             accumulate; the parent tile is written once.
// The translation tables T[octant][theta][phi] are built per level and
// wavelength on the device with the phase formed in FP64 and reduced exactly.
// The root level (the output grid) accumulates in FP64.
#include <algorithm>

#include "wf_common.h"
#include "wf_fft.h"
#include "wf_interp.h"

namespace {

// phi pass: rows of child patterns.  in: [node][F][ntc][npc]; out [node][F][ntc][npp].
__global__ __launch_bounds__(256) void phi_pass_kernel(const float2* __restrict__ in,
                                                       float2* __restrict__ out, int ntc, int npc,
                                                       int npp, int rows_per_block, FftPlanDev pc,
                                                       FftPlanDev pp) {
  extern __shared__ float2 sm[];
  const int lmax = npc > npp ? npc : npp;
  float2* a = sm;
  float2* b = sm + rows_per_block * lmax;
  const int rblocks = (ntc + rows_per_block - 1) / rows_per_block;
  const int64_t plane = blockIdx.x / rblocks;  // node * F + field
  const int row0 = (int)(blockIdx.x - plane * rblocks) * rows_per_block;
  const int nrows = min(rows_per_block, ntc - row0);
  const float2* src = in + (plane * ntc + row0) * (int64_t)npc;
  for (int t = threadIdx.x; t < nrows * npc; t += blockDim.x) {
    const int r = t / npc, k = t - r * npc;
    a[r * lmax + k] = src[(int64_t)r * npc + k];
  }
  __syncthreads();
  float2* X = fft_lines(a, b, nrows, lmax, pc, 0);
  float2* Y = (X == a) ? b : a;
  spectral_resample(X, Y, nrows, lmax, lmax, npc, npp, 0.f, 1.0f / npc);
  __syncthreads();
  float2* Z = fft_lines(Y, X, nrows, lmax, pp, 1);
  float2* dst = out + (plane * ntc + row0) * (int64_t)npp;
  for (int t = threadIdx.x; t < nrows * npp; t += blockDim.x) {
    const int r = t / npp, k = t - r * npp;
    dst[(int64_t)r * npp + k] = Z[r * lmax + k];
  }
}

// theta pass + translation + child sum.
// G: [child node][F][ntc][npp] (phi-interpolated children)
// parent out (float, levels >= 1): [parent][F][ntp][npp]
// root out (double, level 0): [F][ntp][npp]
// The parent tile (W column pairs x 2 ntp) is accumulated in registers: thread
// t owns slots t + 256 m (m < MAXS) of the tile, so no LDS is spent on it.
template <bool ROOT, int MAXS>
__global__ __launch_bounds__(256) void theta_pass_kernel(
    const float2* __restrict__ G, int F, int ntc, int npp, int ntp, int W,
    const int32_t* __restrict__ cb, const int32_t* __restrict__ ce,
    const uint32_t* __restrict__ child_key, const float2* __restrict__ T /*8 x ntp x npp, or null*/,
    float shift, int mirror_base, FftPlanDev pc /*2 ntc*/, FftPlanDev pp /*2 ntp*/,
    float2* __restrict__ out, double2* __restrict__ out_root) {
  extern __shared__ float2 sm[];
  const int n2c = 2 * ntc, n2p = 2 * ntp;
  const int lmax = n2c > n2p ? n2c : n2p;
  float2* a = sm;
  float2* b = sm + W * lmax;
  const int half = npp / 2;
  const int ntiles = (half + W - 1) / W;
  const int tile = (int)(blockIdx.x % ntiles);
  const int64_t pf = blockIdx.x / ntiles;  // parent * F + field
  const int f = (int)(pf % F);
  const int p = (int)(pf / F);
  const int k0 = tile * W;
  const int wcnt = min(W, half - k0);
  const int nslot = wcnt * n2p;
  float2 accf[MAXS];
  double2 accd[ROOT ? MAXS : 1];
#pragma unroll
  for (int m = 0; m < MAXS; ++m) accf[m] = make_float2(0.f, 0.f);
#pragma unroll
  for (int m = 0; m < (ROOT ? MAXS : 1); ++m) accd[m] = make_double2(0.0, 0.0);
  const int c_lo = cb ? cb[p] : p, c_hi = ce ? ce[p] : p + 1;
  for (int c = c_lo; c < c_hi; ++c) {
    __syncthreads();
    const float2* g = G + ((int64_t)c * F + f) * ntc * (int64_t)npp;
    // extended great-circle lines: e[j] = G[j][k], e[ntc + j] = G[ntc-1-j][k + npp/2]
    for (int t = threadIdx.x; t < ntc * wcnt; t += blockDim.x) {
      const int j = t / wcnt, w = t - j * wcnt;
      a[w * lmax + j] = g[(int64_t)j * npp + k0 + w];
      a[w * lmax + ntc + (ntc - 1 - j)] = g[(int64_t)j * npp + k0 + w + half];
    }
    __syncthreads();
    float2* X = fft_lines(a, b, wcnt, lmax, pc, 0);
    float2* Y = (X == a) ? b : a;
    spectral_resample(X, Y, wcnt, lmax, lmax, n2c, n2p, shift, 1.0f / n2c);
    __syncthreads();
    float2* Z = fft_lines(Y, X, wcnt, lmax, pp, 1);
    const int oct = child_key ? (int)(child_key[c] & 7u) : 0;
    const float2* Tt = T ? T + (int64_t)oct * ntp * npp : nullptr;
    // slot s = (w, q): q < ntp -> column k0+w, theta q; else column k0+w+half, theta q-ntp
#pragma unroll
    for (int m = 0; m < MAXS; ++m) {
      const int s = threadIdx.x + 256 * m;
      if (s < nslot) {
        const int w = s / n2p, q = s - w * n2p;
        int i, col, e;
        if (q < ntp) {
          i = q;
          col = k0 + w;
          e = i;
        } else {
          i = q - ntp;
          col = k0 + w + half;
          e = (mirror_base - i) % n2p;
          if (e < 0) e += n2p;
        }
        float2 v = Z[w * lmax + e];
        if (Tt) v = cmul(v, Tt[(int64_t)i * npp + col]);
        if (ROOT) {
          accd[m].x += (double)v.x;
          accd[m].y += (double)v.y;
        } else {
          accf[m] = cadd(accf[m], v);
        }
      }
    }
  }
#pragma unroll
  for (int m = 0; m < MAXS; ++m) {
    const int s = threadIdx.x + 256 * m;
    if (s < nslot) {
      const int w = s / n2p, q = s - w * n2p;
      const int i = q < ntp ? q : q - ntp;
      const int col = q < ntp ? k0 + w : k0 + w + half;
      if (ROOT)
        out_root[((int64_t)f * ntp + i) * npp + col] = accd[m];
      else
        out[(((int64_t)p * F + f) * ntp + i) * npp + col] = accf[m];
    }
  }
}

// Fused level transition for small grids: one block per (parent, field).  For
// each active child: phi pass of all rows into G (kept in LDS), then theta
// pass in batches of column pairs, translation and accumulation into the
// parent held in registers; the parent pattern is written once and every
// child pattern is read once (the compulsory traffic of the level).
template <int MAXS>
__global__ __launch_bounds__(256) void agg_fused_kernel(
    const float2* __restrict__ child, float2* __restrict__ parent, int F, int ntc, int npc,
    int ntp, int npp, int RB, int CB, const int32_t* __restrict__ cb,
    const int32_t* __restrict__ ce, const uint32_t* __restrict__ child_key,
    const float2* __restrict__ T, float shift, int mirror_base, FftPlanDev fpc, FftPlanDev fpp,
    FftPlanDev ftc, FftPlanDev ftp) {
  extern __shared__ float2 sm[];
  const int n2c = 2 * ntc, n2p = 2 * ntp, half = npp / 2;
  const int lphi = npc > npp ? npc : npp;
  const int lth = n2c > n2p ? n2c : n2p;
  const int buf = max(RB * lphi, CB * lth);
  float2* Gs = sm;                      // ntc x npp
  float2* a = sm + ntc * npp;
  float2* b = a + buf;
  const int f = (int)(blockIdx.x % F);
  const int p = (int)(blockIdx.x / F);
  const int npar = ntp * npp;
  float2 acc[MAXS];
#pragma unroll
  for (int m = 0; m < MAXS; ++m) acc[m] = make_float2(0.f, 0.f);
  for (int c = cb[p]; c < ce[p]; ++c) {
    const float2* src = child + ((int64_t)c * F + f) * ntc * (int64_t)npc;
    // phi pass, RB rows at a time
    for (int r0 = 0; r0 < ntc; r0 += RB) {
      const int nr = min(RB, ntc - r0);
      __syncthreads();
      for (int t = threadIdx.x; t < nr * npc; t += blockDim.x) {
        const int r = t / npc, k = t - r * npc;
        a[r * lphi + k] = src[(int64_t)(r0 + r) * npc + k];
      }
      __syncthreads();
      float2* X = fft_lines(a, b, nr, lphi, fpc, 0);
      float2* Y = (X == a) ? b : a;
      spectral_resample(X, Y, nr, lphi, lphi, npc, npp, 0.f, 1.0f / npc);
      __syncthreads();
      float2* Z = fft_lines(Y, X, nr, lphi, fpp, 1);
      for (int t = threadIdx.x; t < nr * npp; t += blockDim.x) {
        const int r = t / npp, k = t - r * npp;
        Gs[(r0 + r) * npp + k] = Z[r * lphi + k];
      }
    }
    const int oct = (int)(child_key[c] & 7u);
    const float2* Tt = T + (int64_t)oct * npar;
    // theta pass, CB column pairs at a time
    for (int k0 = 0; k0 < half; k0 += CB) {
      const int nw = min(CB, half - k0);
      __syncthreads();
      for (int t = threadIdx.x; t < ntc * nw; t += blockDim.x) {
        const int j = t / nw, w = t - j * nw;
        a[w * lth + j] = Gs[j * npp + k0 + w];
        a[w * lth + ntc + (ntc - 1 - j)] = Gs[j * npp + k0 + w + half];
      }
      __syncthreads();
      float2* X = fft_lines(a, b, nw, lth, ftc, 0);
      float2* Y = (X == a) ? b : a;
      spectral_resample(X, Y, nw, lth, lth, n2c, n2p, shift, 1.0f / n2c);
      __syncthreads();
      float2* Z = fft_lines(Y, X, nw, lth, ftp, 1);
      // parent slot s = i * npp + col owned by thread s mod 256
#pragma unroll
      for (int m = 0; m < MAXS; ++m) {
        const int s = threadIdx.x + 256 * m;
        if (s < npar) {
          const int i = s / npp, col = s - i * npp;
          int w = -1, e = 0;
          if (col >= k0 && col < k0 + nw) {
            w = col - k0;
            e = i;
          } else if (col >= k0 + half && col < k0 + half + nw) {
            w = col - k0 - half;
            e = (mirror_base - i) % n2p;
            if (e < 0) e += n2p;
          }
          if (w >= 0) acc[m] = cadd(acc[m], cmul(Z[w * lth + e], Tt[s]));
        }
      }
    }
  }
  float2* dst = parent + ((int64_t)p * F + f) * (int64_t)npar;
#pragma unroll
  for (int m = 0; m < MAXS; ++m) {
    const int s = threadIdx.x + 256 * m;
    if (s < npar) dst[s] = acc[m];
  }
}

// Translation table T[o][i][k] = exp(+i k0 s(theta_i, phi_k).dc_o) with
// dc_o = (sx hx, sy hy, sz hz), signs from the octant code
// (x: o&1 -> +; y: o&2 -> -; z: o&4 -> -), phase in FP64 revolutions.
__global__ __launch_bounds__(256) void trans_table_kernel(float2* __restrict__ T, int ntp, int npp,
                                                          double theta0, double hx, double hy,
                                                          double hz, double k_rev) {
  const int64_t D = (int64_t)ntp * npp;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 8 * D;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int o = (int)(t / D);
    const int64_t d = t - (int64_t)o * D;
    const int i = (int)(d / npp), k = (int)(d - (int64_t)i * npp);
    const double th = theta0 + WF_PI / ntp * i;
    const double ph = WF_TWO_PI / npp * k;
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    const double sx = (o & 1) ? 1.0 : -1.0;
    const double sy = (o & 2) ? -1.0 : 1.0;
    const double sz = (o & 4) ? -1.0 : 1.0;
    const double rev = k_rev * (sx * hx * ct * cp + sy * hy * ct * sp + sz * hz * st);
    T[t] = cis_rev(rev);
  }
}

}  // namespace

namespace wf {

// Dynamic LDS budget per block (bytes) and line elements per batch.
constexpr size_t kLdsBudget = 64 * 1024;
constexpr int kBatchElems = 3584;  // two ping-pong buffers of this many float2 = 56 KB

hipError_t launch_phi_pass(const float2* in, float2* out, int64_t planes, int ntc, int npc, int npp,
                           const FftPlanDev& pc, const FftPlanDev& pp, hipStream_t st) {
  const int lmax = npc > npp ? npc : npp;
  int rows = kBatchElems / lmax;
  if (rows < 1) rows = 1;
  if (rows > ntc) rows = ntc;
  const size_t sh = (size_t)2 * rows * lmax * sizeof(float2);
  if (sh > kLdsBudget) return hipErrorInvalidValue;
  hipFuncSetAttribute((const void*)phi_pass_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                      (int)sh);
  const int64_t nblk = (int64_t)((ntc + rows - 1) / rows) * planes;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  dim3 grid((unsigned)nblk);
  hipLaunchKernelGGL(phi_pass_kernel, grid, dim3(256), sh, st, in, out, ntc, npc, npp, rows, pc,
                     pp);
  return hipGetLastError();
}

constexpr int kThetaSlots = 16;

hipError_t launch_theta_pass(const float2* G, int F, int ntc, int npp, int ntp, int nparents,
                             const int32_t* cb, const int32_t* ce, const uint32_t* child_key,
                             const float2* T, double shift, int mirror_base,
                             const FftPlanDev& pc, const FftPlanDev& pp, float2* out,
                             double2* out_root, hipStream_t st) {
  const int n2c = 2 * ntc, n2p = 2 * ntp;
  const int lmax = n2c > n2p ? n2c : n2p;
  int W = std::min(kBatchElems / lmax, (256 * kThetaSlots) / n2p);
  if (W < 1) W = 1;
  if (W > npp / 2) W = npp / 2 > 0 ? npp / 2 : 1;
  if (W * n2p > 256 * kThetaSlots) return hipErrorInvalidValue;  // 2 ntp above 4096
  const size_t sh = (size_t)2 * W * lmax * sizeof(float2);
  if (sh > kLdsBudget) return hipErrorInvalidValue;
  const int64_t nblk = (int64_t)((npp / 2 + W - 1) / W) * F * nparents;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  dim3 grid((unsigned)nblk);
  if (out_root) {
    hipFuncSetAttribute((const void*)theta_pass_kernel<true, kThetaSlots>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL((theta_pass_kernel<true, kThetaSlots>), grid, dim3(256), sh, st, G, F, ntc,
                       npp, ntp, W, cb, ce, child_key, T, (float)shift, mirror_base, pc, pp, out,
                       out_root);
  } else {
    hipFuncSetAttribute((const void*)theta_pass_kernel<false, kThetaSlots>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL((theta_pass_kernel<false, kThetaSlots>), grid, dim3(256), sh, st, G, F,
                       ntc, npp, ntp, W, cb, ce, child_key, T, (float)shift, mirror_base, pc, pp,
                       out, out_root);
  }
  return hipGetLastError();
}

constexpr int kFusedSlots = 32;

bool fused_fits(int ntc, int npc, int ntp, int npp) {
  if ((int64_t)ntp * npp > 256 * kFusedSlots) return false;
  const int lphi = npc > npp ? npc : npp;
  const int lth = 2 * ntc > 2 * ntp ? 2 * ntc : 2 * ntp;
  const int RB = std::min(ntc, std::max(1, 2048 / lphi));
  const int CB = std::min(npp / 2, std::max(1, 2048 / lth));
  const size_t buf = (size_t)std::max(RB * lphi, CB * lth);
  return ((size_t)ntc * npp + 2 * buf) * sizeof(float2) <= kLdsBudget;
}

hipError_t launch_agg_fused(const float2* child, float2* parent, int F, int ntc, int npc, int ntp,
                            int npp, int nparents, const int32_t* cb, const int32_t* ce,
                            const uint32_t* child_key, const float2* T, double shift,
                            int mirror_base, const FftPlanDev& fpc, const FftPlanDev& fpp,
                            const FftPlanDev& ftc, const FftPlanDev& ftp, hipStream_t st) {
  if (!fused_fits(ntc, npc, ntp, npp)) return hipErrorInvalidValue;
  const int lphi = npc > npp ? npc : npp;
  const int lth = 2 * ntc > 2 * ntp ? 2 * ntc : 2 * ntp;
  const int RB = std::min(ntc, std::max(1, 2048 / lphi));
  const int CB = std::min(npp / 2, std::max(1, 2048 / lth));
  const size_t buf = (size_t)std::max(RB * lphi, CB * lth);
  const size_t sh = ((size_t)ntc * npp + 2 * buf) * sizeof(float2);
  const int64_t nblk = (int64_t)nparents * F;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  hipFuncSetAttribute((const void*)agg_fused_kernel<kFusedSlots>,
                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
  hipLaunchKernelGGL((agg_fused_kernel<kFusedSlots>), dim3((unsigned)nblk), dim3(256), sh, st,
                     child, parent, F, ntc, npc, ntp, npp, RB, CB, cb, ce, child_key, T,
                     (float)shift, mirror_base, fpc, fpp, ftc, ftp);
  return hipGetLastError();
}

hipError_t launch_trans_table(float2* T, int ntp, int npp, double theta0, const double h[3],
                              double k0, hipStream_t st) {
  const int64_t tot = 8LL * ntp * npp;
  int64_t blocks = (tot + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  hipLaunchKernelGGL(trans_table_kernel, dim3((unsigned)blocks), dim3(256), 0, st, T, ntp, npp,
                     theta0, h[0], h[1], h[2], k0 / WF_TWO_PI);
  return hipGetLastError();
}

}  // namespace wf
