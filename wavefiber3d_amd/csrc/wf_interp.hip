// wf_interp.hip -- upward aggregation of the PO tree (rows A9, A11, A12).
//
// Reference per (parent, active child, component): copyandinterp
// (interpolation.h:82-92: fill, copy, 4 cuFFT C2C, 2 (Nt_c + Np_p) row
// copies, Nyquist fix-ups, 2 scalings), translation_compute (fields.h:92-103,
// host FP32 table treePO.cu:511-554) and sumchildren (fields.h:153-170), each
// a separate launch, plus a linear host search for the child
// (treePO.cu:725-732).  Here one level transition is two kernels:
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
template <bool ROOT>
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
  float2* accf = sm + 2 * W * lmax;                          // W x 2 ntp (float)
  double2* accd = reinterpret_cast<double2*>(sm + 2 * W * lmax);  // (double, ROOT)
  const int half = npp / 2;
  const int ntiles = (half + W - 1) / W;
  const int tile = (int)(blockIdx.x % ntiles);
  const int64_t pf = blockIdx.x / ntiles;  // parent * F + field
  const int f = (int)(pf % F);
  const int p = (int)(pf / F);
  const int k0 = tile * W;
  const int wcnt = min(W, half - k0);
  for (int t = threadIdx.x; t < W * n2p; t += blockDim.x) {
    if (ROOT) accd[t] = make_double2(0.0, 0.0);
    else accf[t] = make_float2(0.f, 0.f);
  }
  const int c_lo = cb ? cb[p] : p, c_hi = ce ? ce[p] : p + 1;
  for (int c = c_lo; c < c_hi; ++c) {
    __syncthreads();
    const float2* g = G + ((int64_t)c * F + f) * ntc * (int64_t)npp;
    // extended great-circle lines: e[j] = G[j][k], e[ntc + j] = G[ntc-1-j][k + npp/2]
    for (int t = threadIdx.x; t < ntc * W; t += blockDim.x) {
      const int j = t / W, w = t - j * W;
      if (w < wcnt) {
        a[w * lmax + j] = g[(int64_t)j * npp + k0 + w];
        a[w * lmax + ntc + (ntc - 1 - j)] = g[(int64_t)j * npp + k0 + w + half];
      }
    }
    __syncthreads();
    float2* X = fft_lines(a, b, wcnt, lmax, pc, 0);
    float2* Y = (X == a) ? b : a;
    spectral_resample(X, Y, wcnt, lmax, lmax, n2c, n2p, shift, 1.0f / n2c);
    __syncthreads();
    float2* Z = fft_lines(Y, X, wcnt, lmax, pp, 1);
    const int oct = child_key ? (int)(child_key[c] & 7u) : 0;
    const float2* Tt = T ? T + (int64_t)oct * ntp * npp : nullptr;
    // accumulate: slot (w, i) for column k0+w, slot (w, ntp + i) for column k0+w+half
    for (int t = threadIdx.x; t < wcnt * n2p; t += blockDim.x) {
      const int w = t / n2p, s = t - w * n2p;
      int i, col, e;
      if (s < ntp) {
        i = s;
        col = k0 + w;
        e = i;
      } else {
        i = s - ntp;
        col = k0 + w + half;
        e = mirror_base - i;
        e %= n2p;
        if (e < 0) e += n2p;
      }
      float2 v = Z[w * lmax + e];
      if (Tt) v = cmul(v, Tt[(int64_t)i * npp + col]);
      if (ROOT) {
        accd[w * n2p + s].x += (double)v.x;
        accd[w * n2p + s].y += (double)v.y;
      } else {
        accf[w * n2p + s] = cadd(accf[w * n2p + s], v);
      }
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < wcnt * n2p; t += blockDim.x) {
    const int w = t / n2p, s = t - w * n2p;
    const int i = s < ntp ? s : s - ntp;
    const int col = s < ntp ? k0 + w : k0 + w + half;
    if (ROOT)
      out_root[((int64_t)f * ntp + i) * npp + col] = accd[w * n2p + s];
    else
      out[(((int64_t)p * F + f) * ntp + i) * npp + col] = accf[w * n2p + s];
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

// Dynamic LDS budget per block (bytes).
constexpr size_t kLdsBudget = 64 * 1024;

hipError_t launch_phi_pass(const float2* in, float2* out, int64_t planes, int ntc, int npc, int npp,
                           const FftPlanDev& pc, const FftPlanDev& pp, hipStream_t st) {
  const int lmax = npc > npp ? npc : npp;
  int rows = 8;
  while (rows > 1 && (size_t)2 * rows * lmax * sizeof(float2) > kLdsBudget) rows >>= 1;
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

hipError_t launch_theta_pass(const float2* G, int F, int ntc, int npp, int ntp, int nparents,
                             const int32_t* cb, const int32_t* ce, const uint32_t* child_key,
                             const float2* T, double shift, int mirror_base,
                             const FftPlanDev& pc, const FftPlanDev& pp, float2* out,
                             double2* out_root, hipStream_t st) {
  const int n2c = 2 * ntc, n2p = 2 * ntp;
  const int lmax = n2c > n2p ? n2c : n2p;
  const size_t acc_elem = out_root ? sizeof(double2) : sizeof(float2);
  int W = 16;
  auto lds = [&](int w) { return (size_t)2 * w * lmax * sizeof(float2) + (size_t)w * n2p * acc_elem; };
  while (W > 1 && lds(W) > kLdsBudget) W >>= 1;
  if (W > npp / 2) W = npp / 2 > 0 ? npp / 2 : 1;
  const size_t sh = lds(W);
  if (sh > kLdsBudget) return hipErrorInvalidValue;
  const int64_t nblk = (int64_t)((npp / 2 + W - 1) / W) * F * nparents;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  dim3 grid((unsigned)nblk);
  if (out_root) {
    hipFuncSetAttribute((const void*)theta_pass_kernel<true>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL(theta_pass_kernel<true>, grid, dim3(256), sh, st, G, F, ntc, npp, ntp, W, cb,
                       ce, child_key, T, (float)shift, mirror_base, pc, pp, out, out_root);
  } else {
    hipFuncSetAttribute((const void*)theta_pass_kernel<false>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL(theta_pass_kernel<false>, grid, dim3(256), sh, st, G, F, ntc, npp, ntp, W,
                       cb, ce, child_key, T, (float)shift, mirror_base, pc, pp, out, out_root);
  }
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
