// Host-side check of the analytic normal derivative of the tangent-plane
// currents (current_point_dn) against central differences of the currents
// themselves (current_point), in FP64, over random points, normals and
// incidences, both polarizations, in the lit oblique branch.  Prints the
// worst relative error and exits 0 when it is below 1e-6; a planted error in
// any of the chain-rule terms shows up at the 1e-2 level or above.
#include <cmath>
#include <cstdio>
#include <random>

#include "wf_currents.h"

static WfIncDev make_inc(double th, double ph, char pol, double eta_re, double eta_im, double lam) {
  WfIncDev d{};
  const double ct = std::cos(th), st = std::sin(th), cp = std::cos(ph), sp = std::sin(ph);
  d.w[0] = ct * cp; d.w[1] = ct * sp; d.w[2] = st;
  const double a[3] = {-st * cp, -st * sp, ct}, b[3] = {sp, -cp, 0.0};
  const double b2[3] = {-sp, cp, 0.0};
  for (int c = 0; c < 3; ++c) {
    if (pol == 'M') { d.E[c] = a[c]; d.H[c] = b[c]; } else { d.E[c] = b2[c]; d.H[c] = a[c]; }
  }
  d.k0 = 2 * WF_PI / lam;
  d.z0 = 376.730313668;
  d.sigma = 0.0;
  d.eta_re = eta_re; d.eta_im = eta_im;
  d.ef_re = 0.3; d.ef_im = -1.7; d.hf_re = 2.1; d.hf_im = 0.4;
  d.mode = pol;
  return d;
}

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1.0, 1.0);
  double worst = 0.0;
  int tested = 0;
  for (int trial = 0; trial < 4000; ++trial) {
    const char pol = (trial & 1) ? 'E' : 'M';
    const WfIncDev inc = make_inc(1.4 * U(rng), 3.0 * U(rng), pol, 1.3 + 0.4 * std::fabs(U(rng)),
                                  0.2 * std::fabs(U(rng)), 0.5e-6);
    double p[3] = {2e-6 * U(rng), 2e-6 * U(rng), 2e-6 * U(rng)};
    double n[3] = {U(rng), U(rng), U(rng)};
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int c = 0; c < 3; ++c) n[c] /= nn;
    const double dot = inc.w[0] * n[0] + inc.w[1] * n[1] + inc.w[2] * n[2];
    if (dot > -0.05 || dot < -0.98) continue;  // lit and clear of both branch edges
    for (int j = 0; j < 3; ++j) {
      double v[3] = {0, 0, 0};
      v[j] = 1.0;
      double2 d[6], cp[6], cm[6];
      wfdev::current_point_dn(p, n, 1.5, inc, v, d);
      const double h = 1e-6;
      double np_[3] = {n[0], n[1], n[2]}, nm[3] = {n[0], n[1], n[2]};
      np_[j] += h;
      nm[j] -= h;
      wfdev::current_point(p, np_, 1.5, inc, cp);
      wfdev::current_point(p, nm, 1.5, inc, cm);
      double num = 0, den = 0;
      for (int c = 0; c < 6; ++c) {
        const double fx = (cp[c].x - cm[c].x) / (2 * h), fy = (cp[c].y - cm[c].y) / (2 * h);
        num += (fx - d[c].x) * (fx - d[c].x) + (fy - d[c].y) * (fy - d[c].y);
        den += fx * fx + fy * fy;
      }
      const double e = std::sqrt(num / den);
      if (e > worst) worst = e;
      ++tested;
    }
  }
  // shadowed points: zero derivative
  const WfIncDev inc = make_inc(0.3, 0.2, 'M', 1.55, 0.1, 0.5e-6);
  double p[3] = {1e-6, 0, 0}, n[3] = {inc.w[0], inc.w[1], inc.w[2]}, v[3] = {0, 1, 0};
  double2 d[6];
  wfdev::current_point_dn(p, n, 1.0, inc, v, d);
  for (int c = 0; c < 6; ++c)
    if (d[c].x != 0.0 || d[c].y != 0.0) { std::printf("shadow derivative not zero\n"); return 1; }
  std::printf("tested %d directions, worst relative error %.3e\n", tested, worst);
  return (tested > 1000 && worst < 1e-6) ? 0 : 1;
}
