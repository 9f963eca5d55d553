// wf_currents.h -- per-point tangent-plane Fresnel currents as a
// device function shared by the currents kernel and the adjoint.
// Restates util/tangentplane.h:18-80, util/util.h:365-382 (n0 = 1) and
// util/fields.h:106-150 in FP64.
#pragma once
#include "wf_common.h"

namespace wfdev {

__host__ __device__ inline double2 zsqrt(double2 z) {
  // principal branch, same branch cut as C99 csqrt (sign of the imaginary part)
  if (z.x == 0.0 && z.y == 0.0) return make_double2(0.0, z.y);
  double t = sqrt((fabs(z.x) + hypot(z.x, z.y)) * 0.5);
  if (z.x >= 0.0) return make_double2(t, z.y / (2.0 * t));
  return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}

// util.h:365-368 with complex n: sqrt(1 - (sin t / n)^2)
__host__ __device__ inline double2 cos_internal(double2 n, double st) {
  double2 s = zdiv(make_double2(st, 0.0), n);
  double2 s2 = zmul(s, s);
  return zsqrt(make_double2(1.0 - s2.x, -s2.y));
}

struct Z3 {
  double2 v[3];
};

__host__ __device__ inline Z3 cross_rz(const double a[3], const Z3& b) {
  Z3 r;
  r.v[0] = zsub(zscale(b.v[2], a[1]), zscale(b.v[1], a[2]));
  r.v[1] = zsub(zscale(b.v[0], a[2]), zscale(b.v[2], a[0]));
  r.v[2] = zsub(zscale(b.v[1], a[0]), zscale(b.v[0], a[1]));
  return r;
}
__host__ __device__ inline Z3 cross_zr(const Z3& a, const double b[3]) {
  Z3 r;
  r.v[0] = zsub(zscale(a.v[1], b[2]), zscale(a.v[2], b[1]));
  r.v[1] = zsub(zscale(a.v[2], b[0]), zscale(a.v[0], b[2]));
  r.v[2] = zsub(zscale(a.v[0], b[1]), zscale(a.v[1], b[0]));
  return r;
}

// util/tangentplane.h:18-80 for one point (FP64), then fields.h:106-150:
// cur[0..2] = area E_f J, cur[3..5] = area H_f M.
__host__ __device__ inline void current_point(const double p[3], const double nr[3], double area,
                                              const WfIncDev& inc, double2 cur[6]) {
#pragma clang fp contract(off)
  const double2 zero = make_double2(0.0, 0.0);
  Z3 Ei{{zero, zero, zero}}, Hi{{zero, zero, zero}}, Es, Hs;
  const double dotvalue = inc.w[0] * nr[0] + inc.w[1] * nr[1] + inc.w[2] * nr[2];  // :29
  if (dotvalue <= 0.0) {                                                         // :31
    const double dot = inc.w[0] * p[0] + inc.w[1] * p[1] + inc.w[2] * p[2];      // :32
    double beam = 1.0;
    if (inc.sigma > 0.0) beam = exp(-0.5 * p[2] * p[2] / (inc.sigma * inc.sigma));  // :33-34
    // Ei0 = beam * e^{-i k0 dot}  (:35); FP64 phase with exact reduction
    const double2 e = zcis_rev(-inc.k0 * dot / WF_TWO_PI);
    const double2 Ei0 = zscale(e, beam);
    const double2 Hi0 = zscale(Ei0, 1.0 / inc.z0);  // :36
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Ei.v[c] = zscale(Ei0, inc.E[c]);  // :37
      Hi.v[c] = zscale(Hi0, inc.H[c]);  // :38
    }
    const double2 eta = make_double2(inc.eta_re, inc.eta_im);
    // :40-49 (|n.w| within 1e-10 of 1); |n.w| above 1, possible by rounding for
    // normals slightly longer than 1, is taken as normal incidence too (the
    // oblique branch would divide by |w x n| = 0 or take acos of a value > 1)
    if (fabs(dotvalue) > 1.0 - 1e-10) {
      const double2 one = make_double2(1.0, 0.0);
      const double2 R = (inc.mode == 'M') ? zdiv(zsub(one, eta), zadd(one, eta))
                                          : zdiv(zsub(eta, one), zadd(one, eta));
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Es.v[c] = zmul(Ei.v[c], R);
        Hs.v[c] = zmul(Hi.v[c], R);
      }
    } else {  // :50-63
      double pn[3] = {inc.w[1] * nr[2] - nr[1] * inc.w[2], inc.w[2] * nr[0] - nr[2] * inc.w[0],
                      nr[1] * inc.w[0] - inc.w[1] * nr[0]};
      const double nn = sqrt(pn[0] * pn[0] + pn[1] * pn[1] + pn[2] * pn[2]);
      pn[0] /= nn; pn[1] /= nn; pn[2] /= nn;  // :51
      const double2 factor =
          zadd(zadd(zscale(Ei.v[0], pn[0]), zscale(Ei.v[1], pn[1])), zscale(Ei.v[2], pn[2]));  // :52
      Z3 Eperp, Epar, Hpar, Hperp;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Eperp.v[c] = zscale(factor, pn[c]);     // :53
        Epar.v[c] = zsub(Ei.v[c], Eperp.v[c]);  // :54
      }
      const Z3 t = cross_rz(inc.w, Epar);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Hpar.v[c] = zscale(t.v[c], 1.0 / inc.z0);  // :55
        Hperp.v[c] = zsub(Hi.v[c], Hpar.v[c]);     // :56
      }
      // :58; |n.w| can exceed 1 by rounding for normals of |n| slightly above 1,
      // where acos would return NaN: clamp to normal incidence
      const double lphi = acos(fmin(fabs(dotvalue), 1.0));
      const double st = sin(lphi), ct0 = cos(lphi);
      // rs/rp with n0 = 1 (util.h:370-382): p0 = cos, pt = eta*ct, q0 = cos, qt = ct/eta
      const double2 one = make_double2(1.0, 0.0);
      const double2 c0 = cos_internal(one, st);
      const double2 ct = cos_internal(eta, st);
      (void)ct0;
      const double2 pt = zmul(eta, ct);
      const double2 Rs = zdiv(zsub(c0, pt), zadd(c0, pt));
      const double2 qt = zdiv(ct, eta);
      const double2 Rp = zdiv(zsub(c0, qt), zadd(c0, qt));
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        Es.v[c] = zsub(zmul(Eperp.v[c], Rs), zmul(Epar.v[c], Rp));  // :61
        Hs.v[c] = zsub(zmul(Hpar.v[c], Rp), zmul(Hperp.v[c], Rs));  // :62
      }
    }
  } else {  // :64-67 shadow
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      Es.v[c] = make_double2(-Ei.v[c].x, -Ei.v[c].y);
      Hs.v[c] = make_double2(-Hi.v[c].x, -Hi.v[c].y);
    }
  }
  Z3 Et, Ht;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    Et.v[c] = zadd(Ei.v[c], Es.v[c]);  // :68
    Ht.v[c] = zadd(Hi.v[c], Hs.v[c]);  // :69
  }
  const Z3 M = cross_zr(Et, nr);  // :70
  const Z3 J = cross_rz(nr, Ht);  // :71
  const double2 ef = make_double2(inc.ef_re * area, inc.ef_im * area);
  const double2 hf = make_double2(inc.hf_re * area, inc.hf_im * area);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    cur[c] = zmul(J.v[c], ef);      // fields.h:123-126, :147-150
    cur[3 + c] = zmul(M.v[c], hf);
  }
}

// Derivative of current_point along a change v of the normal (n -> n + e v,
// d/de at e = 0; the position, the area and the incidence fixed), written
// out by the chain rule through util/tangentplane.h:29-71 (lit test, normal
// incidence branch, s/p decomposition q = w x n / |w x n|, Fresnel rs/rp of
// util/util.h:365-382 at the local angle acos|w.n|, reflected fields, J and
// M).  The lit/shadow test and the normal-incidence branch are taken at the
// unperturbed normal, so the result is the derivative inside the branch the
// point is in (one-sided where a branch changes).
__host__ __device__ inline void current_point_dn(const double p[3], const double nr[3], double area,
                                                 const WfIncDev& inc, const double v[3],
                                                 double2 dcur[6]) {
  const double2 zero = make_double2(0.0, 0.0);
  for (int c = 0; c < 6; ++c) dcur[c] = zero;
  const double dotvalue = inc.w[0] * nr[0] + inc.w[1] * nr[1] + inc.w[2] * nr[2];
  if (!(dotvalue <= 0.0)) return;  // shadow: E_tot = H_tot = 0 for any normal
  const double dot = inc.w[0] * p[0] + inc.w[1] * p[1] + inc.w[2] * p[2];
  double beam = 1.0;
  if (inc.sigma > 0.0) beam = exp(-0.5 * p[2] * p[2] / (inc.sigma * inc.sigma));
  const double2 e = zcis_rev(-inc.k0 * dot / WF_TWO_PI);
  const double2 Ei0 = zscale(e, beam);
  const double2 Hi0 = zscale(Ei0, 1.0 / inc.z0);
  Z3 Ei, Hi, Es, Hs, dEs, dHs;
  for (int c = 0; c < 3; ++c) {
    Ei.v[c] = zscale(Ei0, inc.E[c]);
    Hi.v[c] = zscale(Hi0, inc.H[c]);
  }
  const double2 eta = make_double2(inc.eta_re, inc.eta_im);
  const double2 one = make_double2(1.0, 0.0);
  if (fabs(dotvalue) > 1.0 - 1e-10) {  // normal incidence: R fixed
    const double2 R = (inc.mode == 'M') ? zdiv(zsub(one, eta), zadd(one, eta))
                                        : zdiv(zsub(eta, one), zadd(one, eta));
    for (int c = 0; c < 3; ++c) {
      Es.v[c] = zmul(Ei.v[c], R);
      Hs.v[c] = zmul(Hi.v[c], R);
      dEs.v[c] = zero;
      dHs.v[c] = zero;
    }
  } else {
    const double* w = inc.w;
    double pn[3] = {w[1] * nr[2] - nr[1] * w[2], w[2] * nr[0] - nr[2] * w[0],
                    nr[1] * w[0] - w[1] * nr[0]};
    const double dpn[3] = {w[1] * v[2] - v[1] * w[2], w[2] * v[0] - v[2] * w[0],
                           v[1] * w[0] - w[1] * v[0]};
    const double nn = sqrt(pn[0] * pn[0] + pn[1] * pn[1] + pn[2] * pn[2]);
    const double dnn = (pn[0] * dpn[0] + pn[1] * dpn[1] + pn[2] * dpn[2]) / nn;
    double q[3], dq[3];
    for (int c = 0; c < 3; ++c) {
      q[c] = pn[c] / nn;
      dq[c] = (dpn[c] - q[c] * dnn) / nn;
    }
    const double2 factor =
        zadd(zadd(zscale(Ei.v[0], q[0]), zscale(Ei.v[1], q[1])), zscale(Ei.v[2], q[2]));
    const double2 dfactor =
        zadd(zadd(zscale(Ei.v[0], dq[0]), zscale(Ei.v[1], dq[1])), zscale(Ei.v[2], dq[2]));
    Z3 Eperp, Epar, Hpar, Hperp, dEperp, dEpar, dHpar, dHperp;
    for (int c = 0; c < 3; ++c) {
      Eperp.v[c] = zscale(factor, q[c]);
      dEperp.v[c] = zadd(zscale(dfactor, q[c]), zscale(factor, dq[c]));
      Epar.v[c] = zsub(Ei.v[c], Eperp.v[c]);
      dEpar.v[c] = make_double2(-dEperp.v[c].x, -dEperp.v[c].y);
    }
    const Z3 t = cross_rz(w, Epar), dt = cross_rz(w, dEpar);
    for (int c = 0; c < 3; ++c) {
      Hpar.v[c] = zscale(t.v[c], 1.0 / inc.z0);
      dHpar.v[c] = zscale(dt.v[c], 1.0 / inc.z0);
      Hperp.v[c] = zsub(Hi.v[c], Hpar.v[c]);
      dHperp.v[c] = make_double2(-dHpar.v[c].x, -dHpar.v[c].y);
    }
    // local angle: x = |w.n| = -w.n in the lit branch, lphi = acos x
    const double x = fmin(fabs(dotvalue), 1.0);
    const double dx = -(w[0] * v[0] + w[1] * v[1] + w[2] * v[2]);
    const double lphi = acos(x);
    const double st = sin(lphi);
    const double dst = cos(lphi) * (-dx / sqrt(1.0 - x * x));  // d sin(acos x) = -x dx / sqrt(1 - x^2)
    const double2 c0 = cos_internal(one, st);
    const double2 ct = cos_internal(eta, st);
    // c0 = sqrt(1 - st^2): dc0 = -st dst / c0; ct = sqrt(1 - (st/eta)^2): dct = -(st/eta^2) dst / ct
    const double2 dc0 = zdiv(make_double2(-st * dst, 0.0), c0);
    const double2 eta2 = zmul(eta, eta);
    const double2 dct = zdiv(zdiv(make_double2(-st * dst, 0.0), eta2), ct);
    const double2 pt = zmul(eta, ct), dpt = zmul(eta, dct);
    const double2 qt = zdiv(ct, eta), dqt = zdiv(dct, eta);
    const double2 sp = zadd(c0, pt), sq = zadd(c0, qt);
    const double2 Rs = zdiv(zsub(c0, pt), sp);
    const double2 Rp = zdiv(zsub(c0, qt), sq);
    // d (a - b)/(a + b) = 2 (b da - a db) / (a + b)^2
    const double2 dRs = zdiv(zscale(zsub(zmul(pt, dc0), zmul(c0, dpt)), 2.0), zmul(sp, sp));
    const double2 dRp = zdiv(zscale(zsub(zmul(qt, dc0), zmul(c0, dqt)), 2.0), zmul(sq, sq));
    for (int c = 0; c < 3; ++c) {
      Es.v[c] = zsub(zmul(Eperp.v[c], Rs), zmul(Epar.v[c], Rp));
      Hs.v[c] = zsub(zmul(Hpar.v[c], Rp), zmul(Hperp.v[c], Rs));
      dEs.v[c] = zsub(zadd(zmul(dEperp.v[c], Rs), zmul(Eperp.v[c], dRs)),
                      zadd(zmul(dEpar.v[c], Rp), zmul(Epar.v[c], dRp)));
      dHs.v[c] = zsub(zadd(zmul(dHpar.v[c], Rp), zmul(Hpar.v[c], dRp)),
                      zadd(zmul(dHperp.v[c], Rs), zmul(Hperp.v[c], dRs)));
    }
  }
  Z3 Et, Ht;
  for (int c = 0; c < 3; ++c) {
    Et.v[c] = zadd(Ei.v[c], Es.v[c]);
    Ht.v[c] = zadd(Hi.v[c], Hs.v[c]);
  }
  // M = E_t x n, J = n x H_t
  const Z3 dM1 = cross_zr(dEs, nr), dM2 = cross_zr(Et, v);
  const Z3 dJ1 = cross_rz(v, Ht), dJ2 = cross_rz(nr, dHs);
  const double2 ef = make_double2(inc.ef_re * area, inc.ef_im * area);
  const double2 hf = make_double2(inc.hf_re * area, inc.hf_im * area);
  for (int c = 0; c < 3; ++c) {
    dcur[c] = zmul(zadd(dJ1.v[c], dJ2.v[c]), ef);
    dcur[3 + c] = zmul(zadd(dM1.v[c], dM2.v[c]), hf);
  }
}

}  // namespace wfdev
