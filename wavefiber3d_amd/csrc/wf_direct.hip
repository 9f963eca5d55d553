// wf_direct.hip -- depth-0 "plain method" radiation sum in FP64 and
// the far-field intensity epilogue.
//
// Direct sum: util/fields.h:8-38 (current_green_center) summed over all points
// (treePO.cu:652-670 with depth 0, sphere.cu:592-658): for every output
// direction s_d, F[f][d] = sum_n C[n][f] exp(+i k0 s_d.(x_n - c)).  The
// reference launches one transform plus one blocking reduce per direction;
// here one thread owns one direction, points stream through LDS in tiles, and
// all S solves of one wavelength share the phase (F = 6 S complex columns).
// Phase k0 s.(x - c) is formed in FP64 in revolutions and reduced exactly
// before the FP64 sincos.
//
// Finite distance (compareBEM3d/ellipsoid_simple.cu:229-261, current_green):
// the observation point is ob = dist * s about the origin and each point
// contributes C exp(-i k0 |x - ob|) / |x - ob|, relative to the e^{-i k0 D0}
// / D0 (D0 = 1 m) already in the currents: factor (1/r) exp(-i k0 (r - D0)).
// r - dist is formed without cancellation as (|x|^2 - 2 dist s.x) / (r +
// dist) and the common phase k0 (dist - D0) is reduced once on the host, so
// the per-point phase stays accurate at any distance.
#include <algorithm>
#include <cmath>

#include "wf_common.h"

namespace {

constexpr int kDirThreads = 256;
constexpr int kPtTile = 64;

__device__ inline void grid_dir(int64_t d, int nphi, int ntheta, double s[3]) {
  const int j = (int)(d / nphi), kk = (int)(d - (int64_t)j * nphi);
  const double t = -WF_PI / 2 + 1e-10 + WF_PI / ntheta * j;  // treePO.cu:380
  const double p = WF_TWO_PI / nphi * kk;                      // treePO.cu:382
  double st, ct, sp, cp;
  sincos(t, &st, &ct);
  sincos(p, &sp, &cp);
  s[0] = ct * cp;
  s[1] = ct * sp;
  s[2] = st;
}

template <int F, bool FINITE>
__global__ __launch_bounds__(kDirThreads) void direct_fp64_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z, int64_t n,
    const double2* __restrict__ C, double cx, double cy, double cz, double k_rev, int ntheta,
    int nphi, double dist, double rev0, double2* __restrict__ Fout) {
  __shared__ double sr[kPtTile][FINITE ? 4 : 3];
  __shared__ double2 sc[kPtTile][F];
  const int64_t D = (int64_t)ntheta * nphi;
  const int64_t d = (int64_t)blockIdx.x * kDirThreads + threadIdx.x;
  // point range of this block (split over blockIdx.y; partial sums reduced in order)
  const int64_t p0 = n * blockIdx.y / gridDim.y, p1 = n * (blockIdx.y + 1) / gridDim.y;
  Fout += (int64_t)blockIdx.y * F * D;
  double s[3] = {0, 0, 0};
  if (d < D) grid_dir(d, nphi, ntheta, s);
  double2 acc[F];
#pragma unroll
  for (int f = 0; f < F; ++f) acc[f] = make_double2(0.0, 0.0);
  for (int64_t base = p0; base < p1; base += kPtTile) {
    const int cnt = (int)min<int64_t>(kPtTile, p1 - base);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += kDirThreads) {
      const int64_t i = base + t;
      if (FINITE) {
        // absolute coordinates (m) and |x|^2
        const double xa = x[i], ya = y[i], za = z[i];
        sr[t][0] = xa;
        sr[t][1] = ya;
        sr[t][2] = za;
        sr[t][3] = xa * xa + ya * ya + za * za;
      } else {
        // scaled local coordinates: phase in revolutions = s . r
        sr[t][0] = ((double)x[i] - cx) * k_rev;
        sr[t][1] = ((double)y[i] - cy) * k_rev;
        sr[t][2] = ((double)z[i] - cz) * k_rev;
      }
    }
    for (int t = threadIdx.x; t < cnt * F; t += kDirThreads) {
      const int p = t / F, f = t - p * F;
      sc[p][f] = C[(base + p) * F + f];
    }
    __syncthreads();
    for (int p = 0; p < cnt; ++p) {
      double2 g;
      if (FINITE) {
        const double sx = s[0] * sr[p][0] + s[1] * sr[p][1] + s[2] * sr[p][2];
        const double q = sr[p][3] - 2.0 * dist * sx;  // r^2 - dist^2
        const double r = sqrt(dist * dist + q);
        const double dr = q / (r + dist);  // r - dist
        g = zcis_rev(-(dr * k_rev + rev0));
        const double ir = 1.0 / r;
        g.x *= ir;
        g.y *= ir;
      } else {
        const double rev = s[0] * sr[p][0] + s[1] * sr[p][1] + s[2] * sr[p][2];
        g = zcis_rev(rev);
      }
#pragma unroll
      for (int f = 0; f < F; ++f) {
        const double2 c = sc[p][f];
        acc[f].x += c.x * g.x - c.y * g.y;
        acc[f].y += c.x * g.y + c.y * g.x;
      }
    }
  }
  if (d < D) {
#pragma unroll
    for (int f = 0; f < F; ++f) Fout[(int64_t)f * D + d] = acc[f];
  }
}

// Ordered sum of the per-chunk partial patterns: out[f][d] = sum_p part[p][f][d].
__global__ __launch_bounds__(256) void reduce_parts_kernel(const double2* __restrict__ part, int P,
                                                           int64_t FD, double2* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < FD;
       i += (int64_t)gridDim.x * blockDim.x) {
    double2 a = part[i];
    for (int p = 1; p < P; ++p) a = zadd(a, part[(int64_t)p * FD + i]);
    out[i] = a;
  }
}

// util/fields.h:172-204 computeintensity: project E, H onto the plane normal to
// s, E'' = E - Z0 s x H, I = |E''|^2.  F0 is [6 S][D] (field = 6 s + c).
__global__ __launch_bounds__(256) void intensity_kernel(const double2* __restrict__ F0, int S,
                                                        int ntheta, int nphi, double z0,
                                                        float* __restrict__ If,
                                                        float* __restrict__ Ef,
                                                        double* __restrict__ Id,
                                                        double* __restrict__ Ed) {
  const int64_t D = (int64_t)ntheta * nphi;
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int s_ = blockIdx.y;
  if (d >= D) return;
  double s[3];
  grid_dir(d, nphi, ntheta, s);
  const double2* f = F0 + (int64_t)s_ * 6 * D + d;
  double2 E[3] = {f[0], f[D], f[2 * D]};
  double2 H[3] = {f[3 * D], f[4 * D], f[5 * D]};
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
  const double I = o[0].x * o[0].x + o[0].y * o[0].y + o[1].x * o[1].x + o[1].y * o[1].y +
                   o[2].x * o[2].x + o[2].y * o[2].y;
  if (If) If[(int64_t)s_ * D + d] = (float)I;
  if (Id) Id[(int64_t)s_ * D + d] = I;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t e = ((int64_t)s_ * 3 + c) * D + d;
    if (Ef) reinterpret_cast<float2*>(Ef)[e] = make_float2((float)o[c].x, (float)o[c].y);
    if (Ed) reinterpret_cast<double2*>(Ed)[e] = o[c];
  }
}

// The intensity from the transverse components a = t.E'', b = p.E'' of each
// solve (F0 [2 S][D]): I = |a|^2 + |b|^2, E'' = a t + b p with t, p the unit
// vectors of increasing theta and phi at the output direction.
__global__ __launch_bounds__(256) void intensity_transverse_kernel(
    const double2* __restrict__ F0, int S, int ntheta, int nphi, float* __restrict__ If,
    float* __restrict__ Ef, double* __restrict__ Id, double* __restrict__ Ed) {
  const int64_t D = (int64_t)ntheta * nphi;
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int s_ = blockIdx.y;
  if (d >= D) return;
  const double2 a = F0[(int64_t)(2 * s_) * D + d], b = F0[(int64_t)(2 * s_ + 1) * D + d];
  const double I = a.x * a.x + a.y * a.y + b.x * b.x + b.y * b.y;
  if (If) If[(int64_t)s_ * D + d] = (float)I;
  if (Id) Id[(int64_t)s_ * D + d] = I;
  if (Ef || Ed) {
    const int j = (int)(d / nphi), kk = (int)(d - (int64_t)j * nphi);
    const double th = -WF_PI / 2 + 1e-10 + WF_PI / ntheta * j;  // as grid_dir
    const double ph = WF_TWO_PI / nphi * kk;
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    const double t[3] = {-st * cp, -st * sp, ct}, pv[3] = {-sp, cp, 0.0};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double2 o = zadd(zscale(a, t[c]), zscale(b, pv[c]));
      const int64_t e = ((int64_t)s_ * 3 + c) * D + d;
      if (Ef) reinterpret_cast<float2*>(Ef)[e] = make_float2((float)o.x, (float)o.y);
      if (Ed) reinterpret_cast<double2*>(Ed)[e] = o;
    }
  }
}

}  // namespace

namespace wf {

int direct_split(int64_t D) {
  // split the points so the grid holds >= ~2048 blocks (8 per CU)
  const int64_t dblocks = (D + kDirThreads - 1) / kDirThreads;
  return (int)std::min<int64_t>(32, std::max<int64_t>(1, (2048 + dblocks - 1) / dblocks));
}

hipError_t launch_direct_fp64(const float* x, const float* y, const float* z, int64_t n,
                              const double2* C, int F, const double* centre, double k0,
                              int ntheta, int nphi, double2* Fout, double2* scratch,
                              int64_t scratch_elems, hipStream_t st, double dist) {
  const int64_t D = (int64_t)ntheta * nphi;
  const int64_t dblocks = (D + kDirThreads - 1) / kDirThreads;
  int P = direct_split(D);
  while (P > 1 && ((int64_t)P * F * D > scratch_elems || n / P < 256)) --P;
  dim3 grid((unsigned)dblocks, (unsigned)P);
  double2* dst = (P > 1) ? scratch : Fout;
  const double k_rev = k0 / WF_TWO_PI;
  // the currents carry e^{-i k0 D0}/D0 (D0 = 1 m): the common phase
  // k0 (dist - D0) in revolutions, reduced once here so per-point phases
  // keep full precision at any distance
  const double rev0_full = (dist - 1.0) * k_rev;
  const double rev0 = rev0_full - std::nearbyint(rev0_full);
#define WF_DIRECT_CASE(FF)                                                                  \
  case FF:                                                                                  \
    if (dist > 0.0)                                                                         \
      hipLaunchKernelGGL((direct_fp64_kernel<FF, true>), grid, dim3(kDirThreads), 0, st, x, \
                         y, z, n, C, centre[0], centre[1], centre[2], k_rev, ntheta, nphi,  \
                         dist, rev0, dst);                                                  \
    else                                                                                    \
      hipLaunchKernelGGL((direct_fp64_kernel<FF, false>), grid, dim3(kDirThreads), 0, st,   \
                         x, y, z, n, C, centre[0], centre[1], centre[2], k_rev, ntheta,     \
                         nphi, 0.0, 0.0, dst);                                              \
    break;
  switch (F) {
    WF_DIRECT_CASE(6)
    WF_DIRECT_CASE(12)
    WF_DIRECT_CASE(18)
    WF_DIRECT_CASE(24)
    default:
      return hipErrorInvalidValue;
  }
#undef WF_DIRECT_CASE
  if (P > 1) {
    const int64_t FD = (int64_t)F * D;
    hipLaunchKernelGGL(reduce_parts_kernel, dim3((unsigned)std::min<int64_t>(8192, (FD + 255) / 256)),
                       dim3(256), 0, st, scratch, P, FD, Fout);
  }
  return hipGetLastError();
}

hipError_t launch_intensity(const double2* F0, int S, int ntheta, int nphi, float* I, float* E,
                            double* Id, double* Ed, hipStream_t st) {
  const int64_t D = (int64_t)ntheta * nphi;
  const double z0 = sqrt((4e-7 * WF_PI) / 8.8541878e-12);  // util.h:19
  dim3 grid((unsigned)((D + 255) / 256), (unsigned)S);
  hipLaunchKernelGGL(intensity_kernel, grid, dim3(256), 0, st, F0, S, ntheta, nphi, z0, I, E, Id,
                     Ed);
  return hipGetLastError();
}

hipError_t launch_intensity_transverse(const double2* F0, int S, int ntheta, int nphi, float* I,
                                       float* E, double* Id, double* Ed, hipStream_t st) {
  const int64_t D = (int64_t)ntheta * nphi;
  dim3 grid((unsigned)((D + 255) / 256), (unsigned)S);
  hipLaunchKernelGGL(intensity_transverse_kernel, grid, dim3(256), 0, st, F0, S, ntheta, nphi, I,
                     E, Id, Ed);
  return hipGetLastError();
}

}  // namespace wf
