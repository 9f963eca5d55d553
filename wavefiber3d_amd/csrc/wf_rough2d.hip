// wf_rough2d.hip -- PO of a 2D cross-section (the reference's
// compareBEM2d/rough2d.cu:237-585).
//
// The surface is a closed polyline (x_i, y_i), i = 0 .. n-1, traversed
// counter-clockwise.  Sample i sits at vertex i (not the segment midpoint),
// carries the length of segment i -> i+1 (wrapping to 0) and the normal
// (dy, -dx)/|d| of that segment (rough2d.cu:300-306, :489-500).  Per
// incidence:
//
//   currents (rough2d.cu:237-362): plane wave exp(-i k w.x) along
//     w = (cos t cos p, cos t sin p, sin t); TM: E along (-sin t cos p,
//     -sin t sin p, cos t), H = w x E; TE: H along that vector, E = H x w.
//     Lit (w.n <= 0): normal incidence (| |w.n| - 1 | < 1e-15) reflects E and
//     H with r_s(0); otherwise E splits into the part along w x n (r_s) and
//     the rest (r_p via H), reflected about n.  Shadow: total fields 0.
//     J = n x H_t, M = -(n x E_t); stored times the segment length and
//     -w mu0/4 (J) or -w eps0/4 (M) (rough2d.cu:82-111).
//   radiation (rough2d.cu:58-76, :520-560): for each of n_phi azimuths
//     phi_k = 2 pi k / n_phi at the incidence's elevation t, s = (cos t cos
//     phi_k, cos t sin phi_k, sin t), sum over samples of
//     C exp(i k_rho (s_x x + s_y y)), k_rho = k cos t, times the 2D far-field
//     factor sqrt(2i/(pi k_rho D)) exp(-i k_rho D).
//   intensity (rough2d.cu:562-577): E'' = E_t - Z0 s x H_t (t: the parts
//     normal to s), scattering = |E''|^2 D / (2 radius).
//
// Everything is FP64, as in the reference.  The work is small (n up to a few
// thousand samples, a few thousand directions per incidence), so one kernel
// forms the currents of every (sample, incidence) and one kernel gives each
// thread one (direction, incidence) and streams the samples through LDS.
#include <stdint.h>

#include <algorithm>

#include "wf_common.h"
#include "wf_currents.h"
#include "wf_interp.h"

namespace {

using wfdev::Z3;
using wfdev::cross_rz;
using wfdev::cross_zr;
using wfdev::zsqrt;

constexpr int kR2Threads = 256;
constexpr int kR2Tile = 64;

// rough2d.cu:210-215 (n0 = 1): r_s = (cos t - nt c_t)/(cos t + nt c_t),
// c_t = sqrt(1 - (sin t / nt)^2)
__device__ inline double2 r2_rs(double t, double2 nt) {
  const double2 p0 = wfdev::cos_internal(make_double2(1.0, 0.0), sin(t));
  const double2 pt = zmul(nt, wfdev::cos_internal(nt, sin(t)));
  return zdiv(zsub(p0, pt), zadd(p0, pt));
}
// rough2d.cu:217-222 (n0 = 1): r_p = (cos t - c_t/nt)/(cos t + c_t/nt)
__device__ inline double2 r2_rp(double t, double2 nt) {
  const double2 q0 = wfdev::cos_internal(make_double2(1.0, 0.0), sin(t));
  const double2 qt = zdiv(wfdev::cos_internal(nt, sin(t)), nt);
  return zdiv(zsub(q0, qt), zadd(q0, qt));
}

__device__ inline Z3 zscale3(const Z3& a, double2 s) {
  return Z3{{zmul(a.v[0], s), zmul(a.v[1], s), zmul(a.v[2], s)}};
}
__device__ inline Z3 rscale3(const Z3& a, double s) {
  return Z3{{zscale(a.v[0], s), zscale(a.v[1], s), zscale(a.v[2], s)}};
}
__device__ inline Z3 add3(const Z3& a, const Z3& b) {
  return Z3{{zadd(a.v[0], b.v[0]), zadd(a.v[1], b.v[1]), zadd(a.v[2], b.v[2])}};
}

// One thread per (sample i, incidence s): C[s][i][0..5] (J, then M, scaled).
__global__ __launch_bounds__(kR2Threads) void rough2d_currents_kernel(
    const double* __restrict__ X, const double* __restrict__ Y, int n,
    const wf::Rough2dIncDev* __restrict__ incs, double2* __restrict__ C, int* __restrict__ bad) {
  const int i = blockIdx.x * kR2Threads + threadIdx.x;
  const int s = blockIdx.y;
  if (i >= n) return;
  const wf::Rough2dIncDev inc = incs[s];
  const double x = X[i], y = Y[i];
  const int j = (i < n - 1) ? i + 1 : 0;
  const double xn = X[j], yn = Y[j];
  const double len = sqrt((x - xn) * (x - xn) + (y - yn) * (y - yn));  // :486-487
  double2* dst = C + ((int64_t)s * n + i) * 6;
  if (!(len > 0.0) || !isfinite(len)) {  // reported to the caller as an error
    if (s == 0) atomicMin(bad, i);
#pragma unroll
    for (int c = 0; c < 6; ++c) dst[c] = make_double2(0.0, 0.0);
    return;
  }
  const double dx = xn - x, dy = yn - y;
  const double inv = 1.0 / sqrt(dy * dy + dx * dx);
  const double nr[3] = {dy * inv, -dx * inv, 0.0};  // :303-306
  const double2 zero = make_double2(0.0, 0.0);
  const double ph = inc.k * (inc.w[0] * x + inc.w[1] * y);  // :286-287 (z = 0)
  double sp, cp;
  sincos(ph, &sp, &cp);
  const double2 Ei0 = make_double2(cp, -sp);
  const double2 Hi0 = zscale(Ei0, 1.0 / inc.z0);
  Z3 Ei, Hi;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Ei.v[c] = zscale(Ei0, inc.Ed[c]);
    Hi.v[c] = zscale(Hi0, inc.Hd[c]);
  }
  const double dot = inc.w[0] * nr[0] + inc.w[1] * nr[1] + inc.w[2] * nr[2];  // :308
  double2 out[6];
  if (dot <= 0.0) {
    const double2 eta = make_double2(inc.eta_re, inc.eta_im);
    Z3 Es, Hs;
    if (fabs(fabs(dot) - 1.0) < 1e-15) {  // :311-315
      const double2 R = r2_rs(0.0, eta);
      Es = zscale3(Ei, R);
      Hs = zscale3(Hi, R);
    } else {  // :316-342
      // reflection-plane normal w x n, normalised
      double q[3] = {inc.w[1] * nr[2] - inc.w[2] * nr[1], inc.w[2] * nr[0] - inc.w[0] * nr[2],
                     inc.w[0] * nr[1] - inc.w[1] * nr[0]};
      const double qi = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
      q[0] *= qi;
      q[1] *= qi;
      q[2] *= qi;
      const double2 f = zadd(zadd(zscale(Ei.v[0], q[0]), zscale(Ei.v[1], q[1])), zscale(Ei.v[2], q[2]));
      Z3 Eperp, Epar;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Eperp.v[c] = zscale(f, q[c]);
        Epar.v[c] = zsub(Ei.v[c], Eperp.v[c]);
      }
      const double lphi = acos(fabs(dot));
      const double rd[3] = {inc.w[0] - nr[0] * (2.0 * dot), inc.w[1] - nr[1] * (2.0 * dot),
                            inc.w[2] - nr[2] * (2.0 * dot)};
      const Z3 Es_perp = zscale3(Eperp, r2_rs(lphi, eta));
      const Z3 Hs_perp = rscale3(cross_rz(rd, Es_perp), 1.0 / inc.z0);
      const Z3 Hi_par = rscale3(cross_rz(inc.w, Epar), 1.0 / inc.z0);
      const Z3 Hs_par = zscale3(Hi_par, r2_rp(lphi, eta));
      const Z3 Es_par = rscale3(cross_zr(Hs_par, rd), inc.z0);
      Es = add3(Es_perp, Es_par);
      Hs = add3(Hs_perp, Hs_par);
    }
    const Z3 Et = add3(Es, Ei), Ht = add3(Hs, Hi);
    const Z3 M = cross_rz(nr, Et);  // negated below (:348-349)
    const Z3 J = cross_rz(nr, Ht);  // :350
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      out[c] = zscale(J.v[c], inc.fe * len);
      out[3 + c] = zscale(M.v[c], -inc.fh * len);
    }
  } else {
#pragma unroll
    for (int c = 0; c < 6; ++c) out[c] = zero;
  }
#pragma unroll
  for (int c = 0; c < 6; ++c) dst[c] = out[c];
}

__device__ inline void r2_direction(const wf::Rough2dIncDev& inc, int k, int nphi, double dir[3]) {
  const double phi = (double)k / (double)nphi * 2.0 * WF_PI;  // :521
  double sphi, cphi;
  sincos(phi, &sphi, &cphi);
  dir[0] = inc.ct * cphi;
  dir[1] = inc.ct * sphi;
  dir[2] = inc.st;
}

// One thread per (direction k, incidence s, point split z): the six sums over
// samples [n z / P, n (z + 1) / P) into part[s][z][6][nphi].
__global__ __launch_bounds__(kR2Threads) void rough2d_partial_kernel(
    const double* __restrict__ X, const double* __restrict__ Y, int n,
    const wf::Rough2dIncDev* __restrict__ incs, const double2* __restrict__ C, int nphi,
    double2* __restrict__ part) {
  __shared__ double sx[kR2Tile], sy[kR2Tile];
  __shared__ double2 sc[kR2Tile][6];
  const int k = blockIdx.x * kR2Threads + threadIdx.x;
  const int s = blockIdx.y, z = blockIdx.z, P = gridDim.z;
  const wf::Rough2dIncDev inc = incs[s];
  double dir[3];
  r2_direction(inc, k, nphi, dir);
  const double krev = inc.krho / WF_TWO_PI;
  double2 acc[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) acc[c] = make_double2(0.0, 0.0);
  const double2* Cs = C + (int64_t)s * n * 6;
  const int p0 = (int)((int64_t)n * z / P), p1 = (int)((int64_t)n * (z + 1) / P);
  for (int base = p0; base < p1; base += kR2Tile) {
    const int cnt = min(kR2Tile, p1 - base);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += kR2Threads) {
      sx[t] = X[base + t];
      sy[t] = Y[base + t];
    }
    for (int t = threadIdx.x; t < cnt * 6; t += kR2Threads) sc[t / 6][t % 6] = Cs[(int64_t)base * 6 + t];
    __syncthreads();
    for (int p = 0; p < cnt; ++p) {
      const double2 g = zcis_rev(krev * (sx[p] * dir[0] + sy[p] * dir[1]));  // :46-48, :69
#pragma unroll
      for (int c = 0; c < 6; ++c) acc[c] = zadd(acc[c], zmul(sc[p][c], g));
    }
  }
  if (k >= nphi) return;
  double2* dst = part + ((int64_t)s * P + z) * 6 * nphi + k;
#pragma unroll
  for (int c = 0; c < 6; ++c) dst[(int64_t)c * nphi] = acc[c];
}

// One thread per (direction k, incidence s): the splits summed in order, the
// 2D far-field factor, the projection and the intensity.
__global__ __launch_bounds__(kR2Threads) void rough2d_epilogue_kernel(
    const wf::Rough2dIncDev* __restrict__ incs, const double2* __restrict__ part, int P, int nphi,
    double dist, double norm, double* __restrict__ I) {
  const int k = blockIdx.x * kR2Threads + threadIdx.x;
  const int s = blockIdx.y;
  if (k >= nphi) return;
  const wf::Rough2dIncDev inc = incs[s];
  double dir[3];
  r2_direction(inc, k, nphi, dir);
  const double krev = inc.krho / WF_TWO_PI;
  double2 acc[6];
#pragma unroll
  for (int c = 0; c < 6; ++c) acc[c] = make_double2(0.0, 0.0);
  for (int z = 0; z < P; ++z) {
    const double2* src = part + ((int64_t)s * P + z) * 6 * nphi + k;
#pragma unroll
    for (int c = 0; c < 6; ++c) acc[c] = zadd(acc[c], src[(int64_t)c * nphi]);
  }
  // 2D far-field factor sqrt(2i/(pi k_rho D)) exp(-i k_rho D) (:69)
  const double2 G = zmul(zsqrt(make_double2(0.0, 2.0 / (WF_PI * inc.krho * dist))),
                         zcis_rev(-krev * dist));
  Z3 E, H;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    E.v[c] = zmul(acc[c], G);
    H.v[c] = zmul(acc[3 + c], G);
  }
  // :562-574: remove the components along s, then E'' = E - Z0 s x H
  const double2 de = zadd(zadd(zscale(E.v[0], dir[0]), zscale(E.v[1], dir[1])), zscale(E.v[2], dir[2]));
  const double2 dh = zadd(zadd(zscale(H.v[0], dir[0]), zscale(H.v[1], dir[1])), zscale(H.v[2], dir[2]));
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    E.v[c] = zsub(E.v[c], zscale(de, dir[c]));
    H.v[c] = zsub(H.v[c], zscale(dh, dir[c]));
  }
  const Z3 sxh = cross_rz(dir, H);
  double I2 = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double2 e = zsub(E.v[c], zscale(sxh.v[c], inc.z0));
    I2 += e.x * e.x + e.y * e.y;
  }
  I[(int64_t)s * nphi + k] = I2 * norm;  // :577
}

__global__ void r2_set_int_kernel(int* p, int v) { *p = v; }

}  // namespace

namespace wf {

int rough2d_split(int n, int n_inc, int nphi) {
  // enough (direction, incidence, split) blocks to fill the chip, >= 64 samples per split
  const int64_t kb = (nphi + kR2Threads - 1) / kR2Threads;
  int64_t P = (2048 + kb * n_inc - 1) / (kb * n_inc);
  P = std::min<int64_t>(P, 16);
  P = std::min<int64_t>(P, std::max(1, n / 64));
  return (int)std::max<int64_t>(P, 1);
}

hipError_t launch_rough2d(const double* x, const double* y, int n, const Rough2dIncDev* incs,
                          int n_inc, int nphi, double dist, double radius, double2* C,
                          double2* part, int P, double* I, int* bad, hipStream_t st) {
  hipLaunchKernelGGL(r2_set_int_kernel, dim3(1), dim3(1), 0, st, bad, 0x7fffffff);
  const dim3 gc((unsigned)((n + kR2Threads - 1) / kR2Threads), (unsigned)n_inc);
  hipLaunchKernelGGL(rough2d_currents_kernel, gc, dim3(kR2Threads), 0, st, x, y, n, incs, C, bad);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const unsigned kb = (unsigned)((nphi + kR2Threads - 1) / kR2Threads);
  hipLaunchKernelGGL(rough2d_partial_kernel, dim3(kb, (unsigned)n_inc, (unsigned)P), dim3(kR2Threads),
                     0, st, x, y, n, incs, C, nphi, part);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(rough2d_epilogue_kernel, dim3(kb, (unsigned)n_inc), dim3(kR2Threads), 0, st,
                     incs, part, P, nphi, dist, dist / (2.0 * radius), I);
  return hipGetLastError();
}

}  // namespace wf
