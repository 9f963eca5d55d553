// wf_common.h -- shared device helpers and host-side launch declarations for
// the MI355X (gfx950) PO solver.  Internal header; the public C-ABI is
// include/wf_po.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define WF_PI 3.14159265358979323846
#define WF_TWO_PI 6.28318530717958647692

// ------------------------------------------------------------------ device --
__host__ __device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__host__ __device__ inline float2 cconj(float2 a) { return make_float2(a.x, -a.y); }

__host__ __device__ inline double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline double2 zadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__host__ __device__ inline double2 zsub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__host__ __device__ inline double2 zscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
__host__ __device__ inline double2 zdiv(double2 a, double2 b) {
  double d = b.x * b.x + b.y * b.y;
  return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// e^{2 pi i rev} with the argument reduced in FP64: rev - rint(rev) is exact,
// so the FP32 sincos sees |angle| <= pi whatever the magnitude of rev.
__device__ inline float2 cis_rev(double rev) {
  double f = rev - rint(rev);
  float a = (float)(f * WF_TWO_PI);
  float s, c;
  sincosf(a, &s, &c);
  return make_float2(c, s);
}

// FP64 e^{2 pi i rev} with the same exact reduction, then to the nearest
// quarter turn: f = q/4 + r with |r| <= 1/8 (exact, Sterbenz), t = 2 pi r in
// [-pi/4, pi/4], and the Taylor series of sin t to t^17 and cos t to t^16
// (truncation below 1e-17 there), evaluated by Horner in t^2; the quarter
// turn then rotates (cos t, sin t) by i^q.  No branches and no further range
// reduction, unlike a general-purpose sincos.
__host__ __device__ inline double2 zcis_rev(double rev) {
  const double f = rev - rint(rev);
  const double q = rint(4.0 * f);
  const double t = (f - 0.25 * q) * WF_TWO_PI;
  const double t2 = t * t;
  double sp = 1.0 / 355687428096000.0;  // 1/17!
  sp = sp * t2 - 1.0 / 1307674368000.0;
  sp = sp * t2 + 1.0 / 6227020800.0;
  sp = sp * t2 - 1.0 / 39916800.0;
  sp = sp * t2 + 1.0 / 362880.0;
  sp = sp * t2 - 1.0 / 5040.0;
  sp = sp * t2 + 1.0 / 120.0;
  sp = sp * t2 - 1.0 / 6.0;
  const double s = t + t * (t2 * sp);
  double cp = 1.0 / 20922789888000.0;  // 1/16!
  cp = cp * t2 - 1.0 / 87178291200.0;
  cp = cp * t2 + 1.0 / 479001600.0;
  cp = cp * t2 - 1.0 / 3628800.0;
  cp = cp * t2 + 1.0 / 40320.0;
  cp = cp * t2 - 1.0 / 720.0;
  cp = cp * t2 + 1.0 / 24.0;
  cp = cp * t2 - 0.5;
  const double c = 1.0 + t2 * cp;
  switch ((int)q & 3) {
    case 1: return make_double2(-s, c);
    case 2: return make_double2(-c, -s);
    case 3: return make_double2(s, -c);
    default: return make_double2(c, s);
  }
}

// ----------------------------------------------------------- per-solve data --
// Everything the currents kernel needs for one incidence, precomputed on the
// host in FP64 (treePO.cu:501-509, :580-601).
struct WfIncDev {
  double w[3], E[3], H[3];
  double k0, z0, sigma;   // sigma <= 0: no taper
  double eta_re, eta_im;
  double ef_re, ef_im, hf_re, hf_im;
  int mode;               // 'M' or 'E'
  int pad;
};

// Uniform octree geometry (root box and per-level half sizes), FP64.
struct WfBoxDev {
  double lo[3];
  double hi[3];
};

// ----------------------------------------------------------- host launches --
// Each .hip file exposes host launchers; wf_api.hip orchestrates them.
namespace wf {

hipError_t launch_currents(const float* x, const float* y, const float* z, const float* nx,
                           const float* ny, const float* nz, const float* area, int64_t n,
                           const WfIncDev* inc, int n_inc, int fp64, void* out, int64_t stride_n,
                           int64_t stride_s, hipStream_t st);

// Number of point chunks the direct kernel splits into for D directions.
int direct_split(int64_t D);
// scratch: partial sums of the point-split grid (up to 32 x F x D double2;
// fewer splits when smaller, none when null).  dist > 0: observation points
// at distance dist (m) from the origin, exact Green's function (the centre is
// then unused); dist <= 0: far field.
hipError_t launch_direct_fp64(const float* x, const float* y, const float* z, int64_t n,
                              const double2* C, int F, const double* centre, double k0,
                              int ntheta, int nphi, double2* Fout, double2* scratch,
                              int64_t scratch_elems, hipStream_t st, double dist = 0.0);

hipError_t launch_intensity(const double2* F0, int S, int ntheta, int nphi, float* I, float* E,
                            double* Id, double* Ed, hipStream_t st);
// the same from the transverse components (2 per solve, launch_project_transverse)
hipError_t launch_intensity_transverse(const double2* F0, int S, int ntheta, int nphi, float* I,
                                       float* E, double* Id, double* Ed, hipStream_t st);

hipError_t launch_validate(const float* x, const float* y, const float* z, const float* nx,
                           const float* ny, const float* nz, const float* area, int64_t n,
                           float* bbox /*6, device*/, int* bad /*device*/, hipStream_t st);

// Octree: keys + stable radix sort + level lists.
struct OctreeLevels;  // defined in wf_octree.hip consumers via wf_tree.h
}  // namespace wf
