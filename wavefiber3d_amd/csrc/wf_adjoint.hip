// wf_adjoint.hip -- reverse mode of the PO solve (no reference
// counterpart).
//
// The solve is F = A C (currents C -> far-field components F, complex linear
// for fixed geometry) followed by I = |E''|^2.  For a real loss L with
// cotangent dL/dI the reverse pass applies, stage by stage, the Hermitian
// adjoint of each linear map to the cogradient g = dL/dF^*:
//   epilogue:   g_E = (dL/dI) E'',  g_H = (dL/dI) Z0 s x E''
//   root/levels: conj(translation) -> extended great-circle lines ->
//               anterpolation (the adjoint of FFT interpolation: forward FFT on
//               the parent grid, gather of the child band with the conjugate
//               shift, inverse FFT on the child grid, 1/N_child) -> fold the
//               circle back onto the columns phi, phi + pi -> phi anterpolation
//   leaves:     per point Q0[f] = sum_d conj(P[d][n]) g[d][f] and
//               Q_j[f] = sum_d s_dj conj(P[d][n]) g[d][f]
// No forward intermediate is needed by the reverse pass (the maps are linear);
// the forward is recomputed only to get F at the output directions.  Per point
// the gradient is then assembled in FP64:
//   dL/dx = 2 Re sum_f [ i k C_f conj(Q_f) ]                (radiation phase)
//         + 2 Re sum_f [ conj(Q0_f) dC_f/dx ]               (incident phase, taper)
//   dL/dn = 2 Re sum_f [ conj(Q0_f) dC_f/dn ]               (Fresnel currents; analytic
//                                                            derivative, current_point_dn)
//   dL/da = 2 Re sum_f [ conj(Q0_f) C_f / a ]
#include "wf_adjoint.h"
#include "wf_common.h"
#include "wf_currents.h"
#include "wf_fft.h"

namespace {

__device__ inline void out_dir(int64_t d, int nphi, int ntheta, double s[3]) {
  const int j = (int)(d / nphi), kk = (int)(d - (int64_t)j * nphi);
  const double t = -WF_PI / 2 + 1e-10 + WF_PI / ntheta * j;
  const double p = WF_TWO_PI / nphi * kk;
  double st, ct, sp, cp;
  sincos(t, &st, &ct);
  sincos(p, &sp, &cp);
  s[0] = ct * cp;
  s[1] = ct * sp;
  s[2] = st;
}

// dL/dF^* from dL/dI for one solve; also writes the forward intensity.
__global__ __launch_bounds__(256) void epilogue_adj_kernel(const double2* __restrict__ F0,
                                                           const float* __restrict__ gI, int nt,
                                                           int np, double z0,
                                                           double2* __restrict__ gF,
                                                           float* __restrict__ Iout) {
  const int64_t D = (int64_t)nt * np;
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= D) return;
  double s[3];
  out_dir(d, np, nt, s);
  double2 E[3] = {F0[d], F0[D + d], F0[2 * D + d]};
  double2 H[3] = {F0[3 * D + d], F0[4 * D + d], F0[5 * D + d]};
  double2 dv = zadd(zadd(zscale(E[0], s[0]), zscale(E[1], s[1])), zscale(E[2], s[2]));
#pragma unroll
  for (int c = 0; c < 3; ++c) E[c] = zsub(E[c], zscale(dv, s[c]));
  dv = zadd(zadd(zscale(H[0], s[0]), zscale(H[1], s[1])), zscale(H[2], s[2]));
#pragma unroll
  for (int c = 0; c < 3; ++c) H[c] = zsub(H[c], zscale(dv, s[c]));
  double2 o[3];
  o[0] = zsub(E[0], zscale(zsub(zscale(H[2], s[1]), zscale(H[1], s[2])), z0));
  o[1] = zsub(E[1], zscale(zsub(zscale(H[0], s[2]), zscale(H[2], s[0])), z0));
  o[2] = zsub(E[2], zscale(zsub(zscale(H[1], s[0]), zscale(H[0], s[1])), z0));
  if (Iout)
    Iout[d] = (float)(o[0].x * o[0].x + o[0].y * o[0].y + o[1].x * o[1].x + o[1].y * o[1].y +
                      o[2].x * o[2].x + o[2].y * o[2].y);
  const double g = gI[d];
  // s x E''
  const double2 sx = zsub(zscale(o[2], s[1]), zscale(o[1], s[2]));
  const double2 sy = zsub(zscale(o[0], s[2]), zscale(o[2], s[0]));
  const double2 sz = zsub(zscale(o[1], s[0]), zscale(o[0], s[1]));
  gF[d] = zscale(o[0], g);
  gF[D + d] = zscale(o[1], g);
  gF[2 * D + d] = zscale(o[2], g);
  gF[3 * D + d] = zscale(sx, g * z0);
  gF[4 * D + d] = zscale(sy, g * z0);
  gF[5 * D + d] = zscale(sz, g * z0);
}

// Adjoint of theta_pass: parent cogradient -> phi-interpolated child cogradient.
template <bool ROOT>
__global__ __launch_bounds__(256) void theta_adj_kernel(
    const float2* __restrict__ gP, const double2* __restrict__ gProot, int F, int ntc, int npp,
    int ntp, int W, const int32_t* __restrict__ parent_idx, const uint32_t* __restrict__ child_key,
    const float2* __restrict__ T, float shift, int mirror_base, FftPlanDev pp /*2 ntp*/,
    FftPlanDev pc /*2 ntc*/, float2* __restrict__ gG) {
  extern __shared__ float2 sm[];
  const int n2c = 2 * ntc, n2p = 2 * ntp;
  const int lmax = n2c > n2p ? n2c : n2p;
  float2* a = sm;
  float2* b = sm + W * lmax;
  const int half = npp / 2;
  const int ntiles = (half + W - 1) / W;
  const int tile = (int)(blockIdx.x % ntiles);
  const int64_t cf = blockIdx.x / ntiles;  // child * F + field
  const int f = (int)(cf % F);
  const int c = (int)(cf / F);
  const int k0 = tile * W;
  const int wcnt = min(W, half - k0);
  const int p = parent_idx ? parent_idx[c] : 0;
  const int oct = child_key ? (int)(child_key[c] & 7u) : 0;
  const float2* Tt = T ? T + (int64_t)oct * ntp * npp : nullptr;
  const int64_t pbase = ((int64_t)p * F + f) * ntp * (int64_t)npp;
  auto gp = [&](int i, int col) -> float2 {
    float2 v;
    if (ROOT) {
      const double2 dv = gProot[pbase + (int64_t)i * npp + col];
      v = make_float2((float)dv.x, (float)dv.y);
    } else {
      v = gP[pbase + (int64_t)i * npp + col];
    }
    if (Tt) v = cmul(cconj(Tt[(int64_t)i * npp + col]), v);
    return v;
  };
  for (int t = threadIdx.x; t < wcnt * n2p; t += blockDim.x) {
    const int w = t / n2p, e = t - w * n2p;
    float2 v = make_float2(0.f, 0.f);
    if (e < ntp) v = gp(e, k0 + w);
    int i2 = (mirror_base - e) % n2p;
    if (i2 < 0) i2 += n2p;
    if (i2 < ntp) v = cadd(v, gp(i2, k0 + w + half));
    a[w * lmax + e] = v;
  }
  __syncthreads();
  float2* X = fft_lines(a, b, wcnt, lmax, pp, 0);
  float2* Y = (X == a) ? b : a;
  spectral_resample_adj(X, Y, wcnt, lmax, lmax, n2c, n2p, shift, 1.0f / n2c);
  __syncthreads();
  float2* Z = fft_lines(Y, X, wcnt, lmax, pc, 1);
  float2* g = gG + ((int64_t)c * F + f) * ntc * (int64_t)npp;
  for (int t = threadIdx.x; t < ntc * W; t += blockDim.x) {
    const int j = t / W, w = t - j * W;
    if (w < wcnt) {
      g[(int64_t)j * npp + k0 + w] = Z[w * lmax + j];
      g[(int64_t)(ntc - 1 - j) * npp + k0 + w + half] = Z[w * lmax + ntc + j];
    }
  }
}

// Adjoint of phi_pass on rows: [plane][ntc][npp] -> [plane][ntc][npc].
__global__ __launch_bounds__(256) void phi_adj_kernel(const float2* __restrict__ in,
                                                      float2* __restrict__ out, int ntc, int npc,
                                                      int npp, int rows_per_block, FftPlanDev pp,
                                                      FftPlanDev pc) {
  extern __shared__ float2 sm[];
  const int lmax = npc > npp ? npc : npp;
  float2* a = sm;
  float2* b = sm + rows_per_block * lmax;
  const int rblocks = (ntc + rows_per_block - 1) / rows_per_block;
  const int64_t plane = blockIdx.x / rblocks;
  const int row0 = (int)(blockIdx.x - plane * rblocks) * rows_per_block;
  const int nrows = min(rows_per_block, ntc - row0);
  const float2* src = in + (plane * ntc + row0) * (int64_t)npp;
  for (int t = threadIdx.x; t < nrows * npp; t += blockDim.x) {
    const int r = t / npp, k = t - r * npp;
    a[r * lmax + k] = src[(int64_t)r * npp + k];
  }
  __syncthreads();
  float2* X = fft_lines(a, b, nrows, lmax, pp, 0);
  float2* Y = (X == a) ? b : a;
  spectral_resample_adj(X, Y, nrows, lmax, lmax, npc, npp, 0.f, 1.0f / npc);
  __syncthreads();
  float2* Z = fft_lines(Y, X, nrows, lmax, pc, 1);
  float2* dst = out + (plane * ntc + row0) * (int64_t)npc;
  for (int t = threadIdx.x; t < nrows * npc; t += blockDim.x) {
    const int r = t / npc, k = t - r * npc;
    dst[(int64_t)r * npc + k] = Z[r * lmax + k];
  }
}

__device__ inline void leaf_centre(uint32_t key, int depth, const double lo[3],
                                   const double size[3], double c[3]) {
  int ix = 0, iy = 0, iz = 0;
  for (int l = depth - 1; l >= 0; --l) {
    const uint32_t o = (key >> (3 * l)) & 7u;
    ix = 2 * ix + (int)(o & 1u);
    iy = 2 * iy + (int)((o & 2u) == 0);
    iz = 2 * iz + (int)((o & 4u) == 0);
  }
  c[0] = lo[0] + (ix + 0.5) * size[0];
  c[1] = lo[1] + (iy + 0.5) * size[1];
  c[2] = lo[2] + (iz + 0.5) * size[2];
}

constexpr int kAdjTile = 64;
constexpr int kFB = 6;

// Leaf back-projection: Q[n][f][0..3] = sum_d conj(P[d][n]) {1, s_x, s_y, s_z} g[leaf][f][d].
__global__ __launch_bounds__(256) void leaf_adj_kernel(wf::LeafAdjArgs a) {
  __shared__ float3 ss[kAdjTile];
  __shared__ float2 sg[kAdjTile][kFB];
  const int nfb = (a.F + kFB - 1) / kFB;
  const int leaf = (int)(blockIdx.x / nfb);
  const int fb = (int)(blockIdx.x % nfb) * kFB;
  const int D = a.nt * a.np;
  const int b = a.leaf_begin[leaf], e = a.leaf_end[leaf];
  double c[3];
  leaf_centre(a.leaf_key[leaf], a.depth, a.box_lo, a.leaf_size, c);
  const double k_rev = a.k0 / WF_TWO_PI;
  for (int pb = b; pb < e; pb += blockDim.x) {
    const int i = pb + threadIdx.x;
    const bool valid = i < e;
    float r[3] = {0.f, 0.f, 0.f};
    if (valid) {
      r[0] = (float)(((double)a.px[i] - c[0]) * k_rev);
      r[1] = (float)(((double)a.py[i] - c[1]) * k_rev);
      r[2] = (float)(((double)a.pz[i] - c[2]) * k_rev);
    }
    float2 acc[kFB][4];
#pragma unroll
    for (int f = 0; f < kFB; ++f)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[f][j] = make_float2(0.f, 0.f);
    for (int d0 = 0; d0 < D; d0 += kAdjTile) {
      const int cnt = min(kAdjTile, D - d0);
      __syncthreads();
      for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
        const int d = d0 + t;
        const int jj = d / a.np, kk = d - jj * a.np;
        const double th = -WF_PI / 2 + WF_PI * (jj + 0.5) / a.nt;
        const double ph = WF_TWO_PI * kk / a.np;
        double st, ct, sp, cp;
        sincos(th, &st, &ct);
        sincos(ph, &sp, &cp);
        ss[t] = make_float3((float)(ct * cp), (float)(ct * sp), (float)st);
      }
      for (int t = threadIdx.x; t < cnt * kFB; t += blockDim.x) {
        const int dd = t / kFB, f = t - dd * kFB;
        sg[dd][f] = (fb + f < a.F) ? a.g[((int64_t)leaf * a.F + fb + f) * D + d0 + dd]
                                   : make_float2(0.f, 0.f);
      }
      __syncthreads();
      if (valid) {
        for (int dd = 0; dd < cnt; ++dd) {
          const float3 s = ss[dd];
          float rev = s.x * r[0] + s.y * r[1] + s.z * r[2];
          rev -= rintf(rev);
          const float cs = __builtin_amdgcn_cosf(rev), sn = __builtin_amdgcn_sinf(rev);
          // conj(P) = (cs, -sn)
#pragma unroll
          for (int f = 0; f < kFB; ++f) {
            const float2 g = sg[dd][f];
            const float2 v = make_float2(cs * g.x + sn * g.y, cs * g.y - sn * g.x);
            acc[f][0] = cadd(acc[f][0], v);
            acc[f][1] = cadd(acc[f][1], cscale(v, s.x));
            acc[f][2] = cadd(acc[f][2], cscale(v, s.y));
            acc[f][3] = cadd(acc[f][3], cscale(v, s.z));
          }
        }
      }
    }
    if (valid) {
#pragma unroll
      for (int f = 0; f < kFB; ++f) {
        if (fb + f >= a.F) break;
#pragma unroll
        for (int j = 0; j < 4; ++j) a.Q[((int64_t)i * a.F + fb + f) * 4 + j] = acc[f][j];
      }
    }
  }
}

// FP64 direct back-projection over the output grid for one solve (6 fields):
// Q[n][c][j] = sum_d conj(exp(i k s_d.(x_n - centre))) {1, s}_j g[c][d].
__global__ __launch_bounds__(256) void direct_adj_fp64_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    int64_t n, double cx, double cy, double cz, double k_rev, int nt, int np,
    const double2* __restrict__ g /*[6][D]*/, double2* __restrict__ Q /*[n][6][4]*/) {
  __shared__ double3 ss[kAdjTile];
  __shared__ double2 sg[kAdjTile][6];
  const int64_t D = (int64_t)nt * np;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = i < n;
  double r[3] = {0, 0, 0};
  if (valid) {
    r[0] = ((double)px[i] - cx) * k_rev;
    r[1] = ((double)py[i] - cy) * k_rev;
    r[2] = ((double)pz[i] - cz) * k_rev;
  }
  double2 acc[6][4];
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[c][j] = make_double2(0.0, 0.0);
  for (int64_t d0 = 0; d0 < D; d0 += kAdjTile) {
    const int cnt = (int)min<int64_t>(kAdjTile, D - d0);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
      double s[3];
      out_dir(d0 + t, np, nt, s);
      ss[t] = make_double3(s[0], s[1], s[2]);
    }
    for (int t = threadIdx.x; t < cnt * 6; t += blockDim.x) {
      const int dd = t / 6, c = t - dd * 6;
      sg[dd][c] = g[(int64_t)c * D + d0 + dd];
    }
    __syncthreads();
    if (valid) {
      for (int dd = 0; dd < cnt; ++dd) {
        const double3 s = ss[dd];
        const double2 e = zcis_rev(s.x * r[0] + s.y * r[1] + s.z * r[2]);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
          const double2 gg = sg[dd][c];
          const double2 v = make_double2(e.x * gg.x + e.y * gg.y, e.x * gg.y - e.y * gg.x);
          acc[c][0] = zadd(acc[c][0], v);
          acc[c][1] = zadd(acc[c][1], zscale(v, s.x));
          acc[c][2] = zadd(acc[c][2], zscale(v, s.y));
          acc[c][3] = zadd(acc[c][3], zscale(v, s.z));
        }
      }
    }
  }
  if (valid) {
#pragma unroll
    for (int c = 0; c < 6; ++c)
#pragma unroll
      for (int j = 0; j < 4; ++j) Q[(i * 6 + c) * 4 + j] = acc[c][j];
  }
}

template <typename QT>
__device__ inline double2 ld(const QT* q, int64_t i);
template <>
__device__ inline double2 ld<float2>(const float2* q, int64_t i) {
  return make_double2(q[i].x, q[i].y);
}
template <>
__device__ inline double2 ld<double2>(const double2* q, int64_t i) {
  return q[i];
}

// Per-point gradient of one solve (fields f0..f0+5 of Q), accumulated in FP64
// at the sorted position: acc[0..2] dL/dx, acc[3..5] dL/dn, acc[6] dL/da.
template <typename QT>
__global__ __launch_bounds__(256) void grad_assemble_kernel(const float* __restrict__ pts,
                                                            int64_t n, const QT* __restrict__ Q,
                                                            int F, int f0,
                                                            const WfIncDev* __restrict__ incp,
                                                            double* __restrict__ acc) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const WfIncDev inc = *incp;
  const double p[3] = {pts[i], pts[n + i], pts[2 * n + i]};
  const double nr[3] = {pts[3 * n + i], pts[4 * n + i], pts[5 * n + i]};
  const double area = pts[6 * n + i];
  double2 Cu[6];
  wfdev::current_point(p, nr, 1.0, inc, Cu);  // unit-area currents
  double2 q[6][4];
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int j = 0; j < 4; ++j) q[c][j] = ld<QT>(Q, ((i * F) + f0 + c) * 4 + j);
  const double k = inc.k0;
  double gx[3] = {0, 0, 0}, ga = 0, s0re = 0, s0im = 0;
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const double2 C = zscale(Cu[c], area);
    // conj(Q0) C
    const double2 t0 = make_double2(q[c][0].x * C.x + q[c][0].y * C.y, q[c][0].x * C.y - q[c][0].y * C.x);
    s0re += t0.x;
    s0im += t0.y;
    ga += 2.0 * (q[c][0].x * Cu[c].x + q[c][0].y * Cu[c].y);
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      // Im(C conj(Qj)) = C.y Qj.x - C.x Qj.y
      gx[j] += -2.0 * k * (C.y * q[c][j + 1].x - C.x * q[c][j + 1].y);
    }
  }
  // incident phase e^{-i k w.x}: dC/dx_j = -i k w_j C -> 2 Re(conj(Q0)(-i k w_j) C) = 2 k w_j Im(conj(Q0) C)
#pragma unroll
  for (int j = 0; j < 3; ++j) gx[j] += 2.0 * k * inc.w[j] * s0im;
  if (inc.sigma > 0.0) gx[2] += 2.0 * s0re * (-p[2] / (inc.sigma * inc.sigma));
  // normals: the analytic derivative of the Fresnel currents along each axis
  double gn[3] = {0, 0, 0};
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double v[3] = {j == 0 ? 1.0 : 0.0, j == 1 ? 1.0 : 0.0, j == 2 ? 1.0 : 0.0};
    double2 dC[6];
    wfdev::current_point_dn(p, nr, area, inc, v, dC);
    double sacc = 0;
#pragma unroll
    for (int c = 0; c < 6; ++c) sacc += q[c][0].x * dC[c].x + q[c][0].y * dC[c].y;
    gn[j] = 2.0 * sacc;
  }
  acc[i] += gx[0];
  acc[n + i] += gx[1];
  acc[2 * n + i] += gx[2];
  acc[3 * n + i] += gn[0];
  acc[4 * n + i] += gn[1];
  acc[5 * n + i] += gn[2];
  acc[6 * n + i] += ga;
}

__global__ __launch_bounds__(256) void grad_scatter_kernel(const double* __restrict__ acc,
                                                           const int32_t* __restrict__ perm,
                                                           int64_t n, float* __restrict__ gx,
                                                           float* __restrict__ gn,
                                                           float* __restrict__ ga) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t o = perm[i];
    if (gx) {
      gx[o] = (float)acc[i];
      gx[n + o] = (float)acc[n + i];
      gx[2 * n + o] = (float)acc[2 * n + i];
    }
    if (gn) {
      gn[o] = (float)acc[3 * n + i];
      gn[n + o] = (float)acc[4 * n + i];
      gn[2 * n + o] = (float)acc[5 * n + i];
    }
    if (ga) ga[o] = (float)acc[6 * n + i];
  }
}

constexpr size_t kLdsBudget = 64 * 1024;

}  // namespace

namespace wf {

hipError_t launch_epilogue_adj(const double2* F0, const float* gI, int nt, int np, double2* gF,
                               float* Iout, hipStream_t st) {
  const int64_t D = (int64_t)nt * np;
  const double z0 = sqrt((4e-7 * WF_PI) / 8.8541878e-12);
  hipLaunchKernelGGL(epilogue_adj_kernel, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, st, F0,
                     gI, nt, np, z0, gF, Iout);
  return hipGetLastError();
}

hipError_t launch_theta_adj(const float2* gP, const double2* gProot, int F, int ntc, int npp,
                            int ntp, int nchild, const int32_t* parent_idx,
                            const uint32_t* child_key, const float2* T, double shift,
                            int mirror_base, const FftPlanDev& pp, const FftPlanDev& pc,
                            float2* gG, hipStream_t st) {
  const int n2c = 2 * ntc, n2p = 2 * ntp;
  const int lmax = n2c > n2p ? n2c : n2p;
  int W = 16;
  while (W > 1 && (size_t)2 * W * lmax * sizeof(float2) > kLdsBudget) W >>= 1;
  if (W > npp / 2) W = npp / 2 > 0 ? npp / 2 : 1;
  const size_t sh = (size_t)2 * W * lmax * sizeof(float2);
  if (sh > kLdsBudget) return hipErrorInvalidValue;
  const int64_t nblk = (int64_t)((npp / 2 + W - 1) / W) * F * nchild;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  if (gProot) {
    hipFuncSetAttribute((const void*)theta_adj_kernel<true>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL(theta_adj_kernel<true>, dim3((unsigned)nblk), dim3(256), sh, st, gP, gProot,
                       F, ntc, npp, ntp, W, parent_idx, child_key, T, (float)shift, mirror_base,
                       pp, pc, gG);
  } else {
    hipFuncSetAttribute((const void*)theta_adj_kernel<false>,
                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    hipLaunchKernelGGL(theta_adj_kernel<false>, dim3((unsigned)nblk), dim3(256), sh, st, gP,
                       gProot, F, ntc, npp, ntp, W, parent_idx, child_key, T, (float)shift,
                       mirror_base, pp, pc, gG);
  }
  return hipGetLastError();
}

hipError_t launch_phi_adj(const float2* in, float2* out, int64_t planes, int ntc, int npc, int npp,
                          const FftPlanDev& pp, const FftPlanDev& pc, hipStream_t st) {
  const int lmax = npc > npp ? npc : npp;
  int rows = 8;
  while (rows > 1 && (size_t)2 * rows * lmax * sizeof(float2) > kLdsBudget) rows >>= 1;
  if (rows > ntc) rows = ntc;
  const size_t sh = (size_t)2 * rows * lmax * sizeof(float2);
  if (sh > kLdsBudget) return hipErrorInvalidValue;
  const int64_t nblk = (int64_t)((ntc + rows - 1) / rows) * planes;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  hipFuncSetAttribute((const void*)phi_adj_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                      (int)sh);
  hipLaunchKernelGGL(phi_adj_kernel, dim3((unsigned)nblk), dim3(256), sh, st, in, out, ntc, npc,
                     npp, rows, pp, pc);
  return hipGetLastError();
}

hipError_t launch_leaf_adj(const LeafAdjArgs& a, hipStream_t st) {
  if (a.nleaves <= 0) return hipSuccess;
  const int64_t nblk = (int64_t)a.nleaves * ((a.F + kFB - 1) / kFB);
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(leaf_adj_kernel, dim3((unsigned)nblk), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_direct_adj_fp64(const float* pts, int64_t n, const double* centre, double k0,
                                  int nt, int np, const double2* g, double2* Q, hipStream_t st) {
  hipLaunchKernelGGL(direct_adj_fp64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     pts, pts + n, pts + 2 * n, n, centre[0], centre[1], centre[2],
                     k0 / WF_TWO_PI, nt, np, g, Q);
  return hipGetLastError();
}

hipError_t launch_grad_assemble(const float* pts, int64_t n, const void* Q, bool q_fp64, int F,
                                int f0, const WfIncDev* inc, double* acc, hipStream_t st) {
  const dim3 grid((unsigned)((n + 255) / 256));
  if (q_fp64)
    hipLaunchKernelGGL(grad_assemble_kernel<double2>, grid, dim3(256), 0, st, pts, n,
                       (const double2*)Q, F, f0, inc, acc);
  else
    hipLaunchKernelGGL(grad_assemble_kernel<float2>, grid, dim3(256), 0, st, pts, n,
                       (const float2*)Q, F, f0, inc, acc);
  return hipGetLastError();
}

hipError_t launch_grad_scatter(const double* acc, const int32_t* perm, int64_t n, float* gx,
                               float* gn, float* ga, hipStream_t st) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(grad_scatter_kernel, dim3((unsigned)blocks), dim3(256), 0, st, acc, perm, n,
                     gx, gn, ga);
  return hipGetLastError();
}

}  // namespace wf
