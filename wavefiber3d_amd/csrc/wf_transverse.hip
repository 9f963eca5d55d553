// wf_transverse.hip -- the far field's two transverse components from the six
// radiated fields of each solve, on a level grid.
//
// The intensity of a solve is |E''|^2 with E'' = E_perp - z0 s x H_perp
// (util/fields.h:172-204), a vector normal to s.  In the unit vectors of
// increasing theta and phi at s = (cos th cos ph, cos th sin ph, sin th),
//   t = (-sin th cos ph, -sin th sin ph, cos th),  p = (-sin ph, cos ph, 0),
// (s x t = -p, s x p = t) its components are
//   a = t.E'' = t.E - z0 p.H,    b = p.E'' = p.E + z0 t.H,
// linear in the six radiated fields with coefficients of degree one in
// cos/sin th and ph.  Aggregating a and b instead of the six fields carries
// the same far field with a third of the samples: the translation factors
// are scalars, and a, b are band-limited functions of (theta, phi) one
// degree above E and H, inside the sampling margin of every level grid.
// Across a pole t and p change sign, so on the great circle of a column pair
// (phi, phi + pi) the smooth sequences are a and -a, b and -b; the theta
// passes negate the second half of every great circle on input and on
// output (flip).
#include "wf_common.h"
#include "wf_interp.h"

namespace {

__global__ __launch_bounds__(256) void project_transverse_kernel(
    const float2* __restrict__ F, float2* __restrict__ P, int64_t planes /* nodes x S */, int nt,
    int np, float theta0, float dtheta, float dphi, float z0) {
  const int64_t D = (int64_t)nt * np;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= planes * D) return;
  const int64_t ps = idx / D, d = idx - ps * D;  // plane (node, solve), direction
  const int j = (int)(d / np), k = (int)(d - (int64_t)j * np);
  float st, ct, sp, cp;
  sincosf(theta0 + dtheta * j, &st, &ct);
  sincosf(dphi * k, &sp, &cp);
  const float t[3] = {-st * cp, -st * sp, ct}, p[3] = {-sp, cp, 0.f};
  const float2* f = F + ps * 6 * D + d;
  float2 a = make_float2(0.f, 0.f), b = make_float2(0.f, 0.f);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const float2 e = f[c * D], h = f[(3 + c) * D];
    a.x += t[c] * e.x - z0 * p[c] * h.x;
    a.y += t[c] * e.y - z0 * p[c] * h.y;
    b.x += p[c] * e.x + z0 * t[c] * h.x;
    b.y += p[c] * e.y + z0 * t[c] * h.y;
  }
  P[(ps * 2) * D + d] = a;
  P[(ps * 2 + 1) * D + d] = b;
}

}  // namespace

namespace wf {

hipError_t launch_project_transverse(const float2* F, float2* P, int64_t nodes, int S, int nt,
                                     int np, double theta0, double z0, hipStream_t st) {
  const int64_t total = nodes * S * (int64_t)nt * np;
  if (total <= 0) return hipSuccess;
  const int64_t nb = (total + 255) / 256;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(project_transverse_kernel, dim3((unsigned)nb), dim3(256), 0, st, F, P,
                     nodes * S, nt, np, (float)theta0, (float)(WF_PI / nt),
                     (float)(WF_TWO_PI / np), (float)z0);
  return hipGetLastError();
}

}  // namespace wf
