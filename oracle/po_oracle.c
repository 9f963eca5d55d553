/*
 * po_oracle.c -- FP64 CPU restatement of WaveFiber3d's physical-optics path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library (wavefiber3d_amd)
 * links, loads or calls this file; only tests/, the smoke check and the
 * cpu_baseline leg of bench.py use it, as the checker.
 *
 * Each function restates the reference at the cited file:line in double
 * precision.  Pins (closed forms, invariants, brute force) live in
 * tests/test_oracle_*.py.
 */
#include "po_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;
static const double PI = 3.14159265358979323846;

double orc_z0(void) { return sqrt(ORC_MU0 / ORC_EPS0); }

/* --------------------------------------------------------- Fresnel --- */
/* util/util.h:365-368 cos_internal_angle(n, t) = sqrt(1 - sin(t)/n * sin(t)/n)
 * (principal branch) and :370-382 rs/rp with n0 = eta_in = 1
 * (computeLocalTangentPlane is called with comThr(1,0), treePO.cu:604). */
static cplx cos_internal(cplx n, double t) {
  cplx s = sin(t) / n;
  return csqrt(1.0 - s * s);
}

static void fresnel(double t, cplx eta, cplx* rs, cplx* rp) {
  const cplx n0 = 1.0;
  cplx p0 = n0 * cos_internal(n0, t);       /* util.h:372 */
  cplx pt = eta * cos_internal(eta / n0, t); /* util.h:373 */
  *rs = (p0 - pt) / (p0 + pt);               /* util.h:374 */
  cplx q0 = cos_internal(n0, t) / n0;        /* util.h:379 */
  cplx qt = cos_internal(eta, t) / eta;      /* util.h:380 */
  *rp = (q0 - qt) / (q0 + qt);               /* util.h:381 */
}

void orc_fresnel(double theta_loc, double eta_re, double eta_im, double rs[2], double rp[2]) {
  cplx a, b;
  fresnel(theta_loc, eta_re + I * eta_im, &a, &b);
  rs[0] = creal(a); rs[1] = cimag(a);
  rp[0] = creal(b); rp[1] = cimag(b);
}

/* ------------------------------ incidence basis, radiation factors --- */
/* fibersim/treePO.cu:580-601 (and compareMie/sphere.cu:557-579). */
void orc_incidence_basis(double ti, double pi_, char pol, double w[3], double E[3], double H[3]) {
  w[0] = cos(ti) * cos(pi_);
  w[1] = cos(ti) * sin(pi_);
  w[2] = sin(ti);
  double a[3] = {-sin(ti) * cos(pi_), -sin(ti) * sin(pi_), cos(ti)};
  double b[3] = {sin(pi_), -cos(pi_), 0.0};
  double c[3] = {-sin(pi_), cos(pi_), 0.0};
  if (pol == 'M') { /* TM: treePO.cu:587-593 */
    memcpy(E, a, sizeof a);
    memcpy(H, b, sizeof b);
  } else { /* TE: treePO.cu:594-601 */
    memcpy(E, c, sizeof c);
    memcpy(H, a, sizeof a);
  }
}

/* fibersim/treePO.cu:504-509, Dis = 1 m. */
void orc_radiation_factors(double lambda_m, double Ef[2], double Hf[2]) {
  double k0 = 2.0 * PI / lambda_m;
  double omega = k0 * ORC_C0;
  const double D = 1.0;
  cplx g = cexp(-I * k0 * D) / D;
  cplx ef = -I * omega * ORC_MU0 / (4.0 * PI) * g;
  cplx hf = -I * omega * ORC_EPS0 / (4.0 * PI) * g;
  Ef[0] = creal(ef); Ef[1] = cimag(ef);
  Hf[0] = creal(hf); Hf[1] = cimag(hf);
}

/* -------------------------------------------------------- currents --- */
static void cross_rr(const double a[3], const double b[3], double r[3]) {
  /* util.h:338-344 component formula */
  r[0] = a[1] * b[2] - b[1] * a[2];
  r[1] = a[2] * b[0] - b[2] * a[0];
  r[2] = b[1] * a[0] - a[1] * b[0];
}
static void cross_rc(const double a[3], const cplx b[3], cplx r[3]) {
  r[0] = a[1] * b[2] - b[1] * a[2];
  r[1] = a[2] * b[0] - b[2] * a[0];
  r[2] = b[1] * a[0] - a[1] * b[0];
}
static void cross_cr(const cplx a[3], const double b[3], cplx r[3]) {
  r[0] = a[1] * b[2] - b[1] * a[2];
  r[1] = a[2] * b[0] - b[2] * a[0];
  r[2] = b[1] * a[0] - a[1] * b[0];
}

/* util/tangentplane.h:18-80 for one point, then fields.h:106-126 (area) and
 * fields.h:128-150 (E_f on J, H_f on M).  out: 6 complex = 12 doubles. */
static void current_one(const double p[3], const double nrm[3], double area,
                        const double w[3], const double Ed[3], const double Hd[3],
                        double k0, double Z0, double sigma, char mode, cplx eta,
                        cplx Ef, cplx Hf, double* out) {
  cplx Ei[3] = {0, 0, 0}, Hi[3] = {0, 0, 0}, Es[3], Hs[3];
  double dotvalue = w[0] * nrm[0] + w[1] * nrm[1] + w[2] * nrm[2]; /* :29 */
  if (dotvalue <= 0) {                                              /* :31 lit */
    double dot = w[0] * p[0] + w[1] * p[1] + w[2] * p[2];           /* :32 */
    double beam = 1.0;
    if (sigma > 0 && isfinite(sigma)) beam = exp(-0.5 * p[2] * p[2] / (sigma * sigma)); /* :33-34 */
    cplx Ei0 = beam * cexp(-I * k0 * dot);                          /* :35 */
    cplx Hi0 = Ei0 / Z0;                                            /* :36 */
    for (int c = 0; c < 3; ++c) { Ei[c] = Ed[c] * Ei0; Hi[c] = Hd[c] * Hi0; } /* :37-38 */
    const double tol = 1e-10;
    /* :40-49 normal incidence branch; |n.w| > 1 (rounding, |n| slightly above 1)
       is taken as normal incidence too instead of acos/division NaNs */
    if (fabs(dotvalue) > 1.0 - tol) {
      const cplx ein = 1.0;
      cplx R = (mode == 'M') ? (ein - eta) / (ein + eta) : (eta - ein) / (ein + eta);
      for (int c = 0; c < 3; ++c) { Es[c] = Ei[c] * R; Hs[c] = Hi[c] * R; }
    } else { /* :50-63 oblique */
      double pn[3];
      cross_rr(w, nrm, pn);
      double nn = sqrt(pn[0] * pn[0] + pn[1] * pn[1] + pn[2] * pn[2]);
      for (int c = 0; c < 3; ++c) pn[c] /= nn;                      /* :51 */
      cplx factor = Ei[0] * pn[0] + Ei[1] * pn[1] + Ei[2] * pn[2];  /* :52 */
      cplx Eperp[3], Epar[3], Hpar[3], Hperp[3], t[3];
      for (int c = 0; c < 3; ++c) {
        Eperp[c] = pn[c] * factor;                                  /* :53 */
        Epar[c] = Ei[c] - Eperp[c];                                 /* :54 */
      }
      cross_rc(w, Epar, t);
      for (int c = 0; c < 3; ++c) {
        Hpar[c] = t[c] * (1.0 / Z0);                                /* :55 */
        Hperp[c] = Hi[c] - Hpar[c];                                 /* :56 */
      }
      /* :58 (cosine clamped to 1: |n.w| > 1 by rounding would give NaN) */
      double lphi = acos(fmin(fabs(dotvalue), 1.0));
      cplx Rs, Rp;
      fresnel(lphi, eta, &Rs, &Rp);                                 /* :59-60 */
      for (int c = 0; c < 3; ++c) {
        Es[c] = Eperp[c] * Rs - Epar[c] * Rp;                       /* :61 */
        Hs[c] = Hpar[c] * Rp - Hperp[c] * Rs;                       /* :62 */
      }
    }
  } else { /* :64-67 shadow: Ei, Hi stay zero-initialised */
    for (int c = 0; c < 3; ++c) { Es[c] = -Ei[c]; Hs[c] = -Hi[c]; }
  }
  cplx Et[3], Ht[3], M[3], J[3];
  for (int c = 0; c < 3; ++c) { Et[c] = Ei[c] + Es[c]; Ht[c] = Hi[c] + Hs[c]; } /* :68-69 */
  cross_cr(Et, nrm, M);                                             /* :70 */
  cross_rc(nrm, Ht, J);                                             /* :71 */
  for (int c = 0; c < 3; ++c) {
    cplx j = J[c] * area * Ef; /* fields.h:123-126 then :147-150 */
    cplx m = M[c] * area * Hf;
    out[2 * c] = creal(j);
    out[2 * c + 1] = cimag(j);
    out[6 + 2 * c] = creal(m);
    out[6 + 2 * c + 1] = cimag(m);
  }
}

void orc_currents(int64_t n, const float* x, const float* y, const float* z,
                  const float* nx, const float* ny, const float* nz, const float* area,
                  const orc_incidence* inc, double* C) {
  double w[3], Ed[3], Hd[3], ef[2], hf[2];
  orc_incidence_basis(inc->theta, inc->phi, inc->pol, w, Ed, Hd);
  orc_radiation_factors(inc->lambda_m, ef, hf);
  double k0 = 2.0 * PI / inc->lambda_m;
  double Z0 = orc_z0();
  cplx eta = inc->eta_re + I * inc->eta_im;
  cplx Ef = ef[0] + I * ef[1], Hf = hf[0] + I * hf[1];
  for (int64_t i = 0; i < n; ++i) {
    double p[3] = {x[i], y[i], z[i]};
    double nr[3] = {nx[i], ny[i], nz[i]};
    current_one(p, nr, area[i], w, Ed, Hd, k0, Z0, inc->beam_sigma_m, inc->pol, eta, Ef, Hf,
                C + 12 * i);
  }
}

/* ----------------------------------------------------- output grid --- */
void orc_output_dirs(int ntheta, int nphi, double* dirs) {
  for (int j = 0; j < ntheta; ++j) {
    double t = -PI / 2 + 1e-10 + PI / ntheta * j; /* treePO.cu:380 */
    for (int k = 0; k < nphi; ++k) {
      double p = 2 * PI / nphi * k;               /* treePO.cu:382 */
      double* d = dirs + 3 * ((int64_t)j * nphi + k);
      d[0] = cos(t) * cos(p);
      d[1] = cos(t) * sin(p);
      d[2] = sin(t);
    }
  }
}

/* ------------------------------------------------------ direct sum --- */
void orc_direct_sum(int64_t n, const float* x, const float* y, const float* z,
                    const double centre[3], double k, int S, const double* C,
                    int64_t ndir, const double* dirs, double* F, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    cplx* acc = (cplx*)malloc(sizeof(cplx) * 6 * (size_t)S);
#pragma omp for schedule(dynamic, 16)
    for (int64_t d = 0; d < ndir; ++d) {
      const double* s = dirs + 3 * d;
      for (int q = 0; q < 6 * S; ++q) acc[q] = 0;
      for (int64_t i = 0; i < n; ++i) {
        double dx = (double)x[i] - centre[0];
        double dy = (double)y[i] - centre[1];
        double dz = (double)z[i] - centre[2];
        double ph = k * (dx * s[0] + dy * s[1] + dz * s[2]); /* fields.h:105-109 */
        cplx g = cos(ph) + I * sin(ph);
        for (int sv = 0; sv < S; ++sv) {
          const double* c = C + ((size_t)sv * n + i) * 12;
          cplx* a = acc + 6 * sv;
          for (int q = 0; q < 6; ++q) a[q] += (c[2 * q] + I * c[2 * q + 1]) * g;
        }
      }
      for (int sv = 0; sv < S; ++sv)
        for (int q = 0; q < 6; ++q) {
          double* f = F + (((size_t)sv * ndir + d) * 6 + q) * 2;
          f[0] = creal(acc[6 * sv + q]);
          f[1] = cimag(acc[6 * sv + q]);
        }
    }
    free(acc);
  }
}

/* Finite-distance sum (compareBEM3d/ellipsoid_simple.cu:229-261 and its
 * observation loop :547-555): ob = dist * s_d, green = exp(-i k |x - ob|)/|x - ob|
 * with the distance formed from the coordinate differences as there.  The
 * currents C carry the far-field factor exp(-i k D0)/D0 with D0 = 1 m
 * (treePO.cu:504-509); the reference's validator applies -i w mu0/(4 pi) alone,
 * so the sum is multiplied by D0 exp(+i k D0) to undo it. */
void orc_direct_sum_finite(int64_t n, const float* x, const float* y, const float* z, double k,
                           int S, const double* C, int64_t ndir, const double* dirs, double dist,
                           double* F, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const cplx undo = cos(k) + I * sin(k);
#pragma omp parallel
  {
    cplx* acc = (cplx*)malloc(sizeof(cplx) * 6 * (size_t)S);
#pragma omp for schedule(dynamic, 16)
    for (int64_t d = 0; d < ndir; ++d) {
      const double* s = dirs + 3 * d;
      const double ob[3] = {dist * s[0], dist * s[1], dist * s[2]};
      for (int q = 0; q < 6 * S; ++q) acc[q] = 0;
      for (int64_t i = 0; i < n; ++i) {
        double dx = (double)x[i] - ob[0];
        double dy = (double)y[i] - ob[1];
        double dz = (double)z[i] - ob[2];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        cplx g = (cos(k * r) - I * sin(k * r)) / r;
        for (int sv = 0; sv < S; ++sv) {
          const double* c = C + ((size_t)sv * n + i) * 12;
          cplx* a = acc + 6 * sv;
          for (int q = 0; q < 6; ++q) a[q] += (c[2 * q] + I * c[2 * q + 1]) * g;
        }
      }
      for (int sv = 0; sv < S; ++sv)
        for (int q = 0; q < 6; ++q) {
          cplx v = acc[6 * sv + q] * undo;
          double* f = F + (((size_t)sv * ndir + d) * 6 + q) * 2;
          f[0] = creal(v);
          f[1] = cimag(v);
        }
    }
    free(acc);
  }
}

/* ------------------------------------------------------- intensity --- */
void orc_intensity(int64_t ndir, const double* dirs, const double* F, double* Iout, double* E) {
  const double Z0 = orc_z0();
  for (int64_t d = 0; d < ndir; ++d) {
    const double* s = dirs + 3 * d;
    const double* f = F + 12 * d;
    cplx Ex = f[0] + I * f[1], Ey = f[2] + I * f[3], Ez = f[4] + I * f[5];
    cplx Hx = f[6] + I * f[7], Hy = f[8] + I * f[9], Hz = f[10] + I * f[11];
    cplx dv = Ex * s[0] + Ey * s[1] + Ez * s[2]; /* fields.h:181-184 */
    Ex -= dv * s[0]; Ey -= dv * s[1]; Ez -= dv * s[2];
    dv = Hx * s[0] + Hy * s[1] + Hz * s[2];      /* fields.h:185-188 */
    Hx -= dv * s[0]; Hy -= dv * s[1]; Hz -= dv * s[2];
    Ex = Ex - (Hz * s[1] - Hy * s[2]) * Z0;      /* fields.h:190-192 */
    Ey = Ey - (-(Hz * s[0] - Hx * s[2])) * Z0;
    Ez = Ez - (Hy * s[0] - Hx * s[1]) * Z0;
    Iout[d] = creal(Ex) * creal(Ex) + cimag(Ex) * cimag(Ex) + creal(Ey) * creal(Ey) +
              cimag(Ey) * cimag(Ey) + creal(Ez) * creal(Ez) + cimag(Ez) * cimag(Ez); /* :194 */
    if (E) {
      E[2 * d] = creal(Ex); E[2 * d + 1] = cimag(Ex);
      E[2 * (ndir + d)] = creal(Ey); E[2 * (ndir + d) + 1] = cimag(Ey);
      E[2 * (2 * ndir + d)] = creal(Ez); E[2 * (2 * ndir + d) + 1] = cimag(Ez);
    }
  }
}

/* ----------------------------------------------------- full depth-0 ----- */
static int solve_dirs(int64_t n, const float* x, const float* y, const float* z,
                      const float* nx, const float* ny, const float* nz, const float* area,
                      const orc_incidence* inc, int n_inc, int64_t ndir, const double* dirs,
                      const double centre[3], double dist, double* Iout, double* E,
                      int nthreads) {
  double* C = (double*)malloc(sizeof(double) * 12 * (size_t)n);
  double* F = (double*)malloc(sizeof(double) * 12 * (size_t)ndir);
  if (!C || !F) { free(C); free(F); return 1; }
  for (int i = 0; i < n_inc; ++i) {
    orc_currents(n, x, y, z, nx, ny, nz, area, inc + i, C);
    double k0 = 2.0 * PI / inc[i].lambda_m;
    if (dist > 0)
      orc_direct_sum_finite(n, x, y, z, k0, 1, C, ndir, dirs, dist, F, nthreads);
    else
      orc_direct_sum(n, x, y, z, centre, k0, 1, C, ndir, dirs, F, nthreads);
    orc_intensity(ndir, dirs, F, Iout + (size_t)i * ndir, E ? E + (size_t)i * 6 * ndir : NULL);
  }
  free(C);
  free(F);
  return 0;
}

int orc_solve_direct_dirs(int64_t n, const float* x, const float* y, const float* z,
                          const float* nx, const float* ny, const float* nz, const float* area,
                          const orc_incidence* inc, int n_inc, int64_t ndir, const double* dirs,
                          const double centre[3], double* Iout, double* E, int nthreads) {
  return solve_dirs(n, x, y, z, nx, ny, nz, area, inc, n_inc, ndir, dirs, centre, 0.0, Iout, E,
                    nthreads);
}

int orc_solve_finite_dirs(int64_t n, const float* x, const float* y, const float* z,
                          const float* nx, const float* ny, const float* nz, const float* area,
                          const orc_incidence* inc, int n_inc, int64_t ndir, const double* dirs,
                          double dist, double* Iout, double* E, int nthreads) {
  const double origin[3] = {0.0, 0.0, 0.0};
  return solve_dirs(n, x, y, z, nx, ny, nz, area, inc, n_inc, ndir, dirs, origin, dist, Iout, E,
                    nthreads);
}

int orc_solve_direct(int64_t n, const float* x, const float* y, const float* z,
                     const float* nx, const float* ny, const float* nz, const float* area,
                     const orc_incidence* inc, int n_inc, int ntheta, int nphi,
                     const double centre[3], double* Iout, double* E, int nthreads) {
  int64_t ndir = (int64_t)ntheta * nphi;
  double* dirs = (double*)malloc(sizeof(double) * 3 * (size_t)ndir);
  if (!dirs) return 1;
  orc_output_dirs(ntheta, nphi, dirs);
  int rc = orc_solve_direct_dirs(n, x, y, z, nx, ny, nz, area, inc, n_inc, ndir, dirs, centre,
                                 Iout, E, nthreads);
  free(dirs);
  return rc;
}

/* ---------------------------------------------------------- octree --- */
/* util/tree.h:289-715 restated: node = nodes[level offset + id]; children of id
 * m are 8m..8m+7 at the next level (tree.h:647-662); the centre is
 * 0.5f*(min+max) in float (tree.h:116-120); octants in the order of the eight
 * predicates at tree.h:379-414 with child boxes tree.h:672-687; the partition
 * is stable (warp-ordered prefix at tree.h:492-625, warps own consecutive
 * chunks); there is no minimum-size stop (tree.h:327-330). */
static int octant_of(float px, float py, float pz, float cx, float cy, float cz) {
  if (px < cx && py >= cy && pz >= cz) return 0;
  if (px >= cx && py >= cy && pz >= cz) return 1;
  if (px < cx && py < cy && pz >= cz) return 2;
  if (px >= cx && py < cy && pz >= cz) return 3;
  if (px < cx && py >= cy && pz < cz) return 4;
  if (px >= cx && py >= cy && pz < cz) return 5;
  if (px < cx && py < cy && pz < cz) return 6;
  if (px >= cx && py < cy && pz < cz) return 7;
  return -1; /* NaN coordinates match no predicate and are dropped by the reference */
}

static void octree_rec(int level, int depth, int64_t id, int32_t b, int32_t e, const float* box,
                       const float* x, const float* y, const float* z, int64_t* perm,
                       int64_t* tmp, int32_t* nb, int32_t* ne, float* nbox) {
  int64_t off = 0, p8 = 1;
  for (int l = 0; l < level; ++l) { off += p8; p8 *= 8; }
  int64_t idx = off + id;
  nb[idx] = b;
  ne[idx] = e;
  memcpy(nbox + 6 * idx, box, 6 * sizeof(float));
  if (level >= depth) return;
  float cx = 0.5f * (box[0] + box[3]);
  float cy = 0.5f * (box[1] + box[4]);
  float cz = 0.5f * (box[2] + box[5]);
  int32_t cnt[8] = {0};
  for (int32_t i = b; i < e; ++i) {
    int64_t q = perm[i];
    int o = octant_of(x[q], y[q], z[q], cx, cy, cz);
    if (o >= 0) cnt[o]++;
  }
  int32_t start[9];
  start[0] = b;
  for (int o = 0; o < 8; ++o) start[o + 1] = start[o] + cnt[o];
  int32_t fill[8];
  for (int o = 0; o < 8; ++o) fill[o] = start[o];
  for (int32_t i = b; i < e; ++i) {
    int64_t q = perm[i];
    int o = octant_of(x[q], y[q], z[q], cx, cy, cz);
    if (o >= 0) tmp[fill[o]++] = q;
  }
  for (int32_t i = b; i < start[8]; ++i) perm[i] = tmp[i];
  /* child boxes: tree.h:672-687 */
  float cb[8][6] = {
      {box[0], cy, cz, cx, box[4], box[5]}, {cx, cy, cz, box[3], box[4], box[5]},
      {box[0], box[1], cz, cx, cy, box[5]}, {cx, box[1], cz, box[3], cy, box[5]},
      {box[0], cy, box[2], cx, box[4], cz}, {cx, cy, box[2], box[3], box[4], cz},
      {box[0], box[1], box[2], cx, cy, cz}, {cx, box[1], box[2], box[3], cy, cz}};
  for (int o = 0; o < 8; ++o)
    octree_rec(level + 1, depth, 8 * id + o, start[o], start[o + 1], cb[o], x, y, z, perm, tmp,
               nb, ne, nbox);
}

int orc_octree(int64_t n, const float* x, const float* y, const float* z, const float box[6],
               int depth, int64_t* perm, int32_t* node_begin, int32_t* node_end,
               float* node_box) {
  if (depth < 0 || depth > 8 || n < 0 || n > 0x7fffffff) return 1;
  int64_t* tmp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  if (!tmp) return 1;
  for (int64_t i = 0; i < n; ++i) perm[i] = i;
  octree_rec(0, depth, 0, 0, (int32_t)n, box, x, y, z, perm, tmp, node_begin, node_end, node_box);
  free(tmp);
  return 0;
}

/* -------------------------------------------------- level sampling --- */
int orc_ref_level_count(int level, float extent, float beta) {
  /* treePO.cu:346-351: float d = pow(0.5, i) * extent; ceil(2 pi/0.4e-6 d + beta (..)^(1/3)) */
  float d = (float)(pow(0.5, level) * extent);
  double kd = 2 * PI / (0.4e-6) * d;
  return (int)ceil(kd + beta * pow(kd, 1.0 / 3.0));
}

/* --------------------------------------------------- interpolation --- */
/* Coefficients of the trigonometric interpolant (interpolation.h:17-37 and
 * :47-68: forward DFT, spectral zero-pad, Nyquist split when n is even). */
static void trig_coeffs(int nc, const double* in, cplx* c) {
  for (int m = 0; m < nc; ++m) {
    cplx s = 0;
    for (int j = 0; j < nc; ++j) {
      double a = -2 * PI * (double)((int64_t)m * j % nc) / nc;
      s += (in[2 * j] + I * in[2 * j + 1]) * (cos(a) + I * sin(a));
    }
    c[m] = s / nc; /* interpolation.h:45,77 scale by 1/N_initial */
  }
}

static cplx trig_eval(int nc, const cplx* c, double tau /* t - off_c */) {
  cplx s = 0;
  int half = nc / 2;
  for (int m = 0; m < nc; ++m) {
    int f = (m <= (nc - 1) / 2) ? m : m - nc; /* signed frequency */
    if ((nc % 2 == 0) && m == half) {
      /* setvalue(): Nyquist bin halved and mirrored (interpolation.h:2-9) */
      s += 0.5 * c[m] * (cexp(I * half * tau) + cexp(-I * half * tau));
      continue;
    }
    s += c[m] * cexp(I * f * tau);
  }
  return s;
}

void orc_trig_interp(int nc, const double* in, double off_c, int np, double off_p, double* out) {
  cplx* c = (cplx*)malloc(sizeof(cplx) * nc);
  trig_coeffs(nc, in, c);
  for (int m = 0; m < np; ++m) {
    double t = off_p + 2 * PI * m / np;
    cplx v = trig_eval(nc, c, t - off_c);
    out[2 * m] = creal(v);
    out[2 * m + 1] = cimag(v);
  }
  free(c);
}

void orc_sphere_interp(int ntc, int npc, double off_tc, const double* in, int ntp, int npp,
                       double off_tp, double* out) {
  (void)off_tc; /* the child grid must be the half-sample grid off_tc = -pi/2 + pi/(2 ntc) */
  /* phi pass: each child row to the parent phi grid (same origin phi = 0). */
  double* g = (double*)malloc(sizeof(double) * 2 * (size_t)ntc * npp);
  for (int j = 0; j < ntc; ++j)
    orc_trig_interp(npc, in + 2 * (size_t)j * npc, 0.0, npp, 0.0, g + 2 * (size_t)j * npp);
  /* theta pass on the great circle through columns k and k + npp/2:
   * circle angle a = theta on column k, a = pi - theta on column k + npp/2.
   * With the half-sample child grid the 2 ntc circle samples are uniform,
   * a_j = off + pi j/ntc, off = -pi/2 + pi/(2 ntc). */
  double off = -PI / 2 + PI / (2.0 * ntc);
  int n2 = 2 * ntc;
  double* line = (double*)malloc(sizeof(double) * 2 * n2);
  cplx* c = (cplx*)malloc(sizeof(cplx) * n2);
  for (int k = 0; k < npp / 2; ++k) {
    int k2 = k + npp / 2;
    for (int j = 0; j < ntc; ++j) {
      line[2 * j] = g[2 * ((size_t)j * npp + k)];
      line[2 * j + 1] = g[2 * ((size_t)j * npp + k) + 1];
      int jj = ntc - 1 - j; /* a_{ntc+j} = pi - theta_{ntc-1-j} */
      line[2 * (ntc + j)] = g[2 * ((size_t)jj * npp + k2)];
      line[2 * (ntc + j) + 1] = g[2 * ((size_t)jj * npp + k2) + 1];
    }
    trig_coeffs(n2, line, c);
    for (int i = 0; i < ntp; ++i) {
      double th = off_tp + PI * i / ntp;
      cplx v1 = trig_eval(n2, c, th - off);
      cplx v2 = trig_eval(n2, c, (PI - th) - off);
      out[2 * ((size_t)i * npp + k)] = creal(v1);
      out[2 * ((size_t)i * npp + k) + 1] = cimag(v1);
      out[2 * ((size_t)i * npp + k2)] = creal(v2);
      out[2 * ((size_t)i * npp + k2) + 1] = cimag(v2);
    }
  }
  free(c);
  free(line);
  free(g);
}

/* ---------------------------------------------------------------------------
 * Geometry front-end: util/geometry.h:78-178 (set_sourcepoint with the arc
 * window [start, end]; util/geometry.h:3-72 is start = 0, end = numel - 1).
 * Float arithmetic as there: Float3 components are float, subtract / add /
 * crossproduct (util/util.h:339-345) / vecnorm (util.h:306-309) / normalize
 * (util.h:312-320) work component-wise in float, sums run left to right, the
 * ring height i*unitz - period*0.5 and the /4.0 and 0.5* factors go through
 * double; the build disables contraction.  Returns 0, or 1 + i*numel + j of
 * the first sample in (i, j) loop order with a vertex at x = y = 0 (the
 * reference prints it and exits). */
long orc_heightfield_points(const float* h, int numel, int numelvert, float period, float unitz,
                            int start, int end, float* x, float* y, float* z, float* nx,
                            float* ny, float* nz, float* area) {
  const float* hy = h + (size_t)numel * numelvert;
  const int W = end - start + 1;
  for (int i = 0; i < numelvert; ++i) {
    const float zi = (float)(i * unitz - period * 0.5);
    const float zn = (float)((i + 1) * unitz - period * 0.5);
    for (int j = 0; j < numel; ++j) {
      if (j < start || j > end) continue;
      const int jp = (j < numel - 1) ? j + 1 : 0;
      const int ip = (i < numelvert - 1) ? i + 1 : i;
      const float p1[3] = {h[(size_t)j * numelvert + i], hy[(size_t)j * numelvert + i], zi};
      const float p2[3] = {h[(size_t)jp * numelvert + i], hy[(size_t)jp * numelvert + i], zi};
      const float p3[3] = {h[(size_t)j * numelvert + ip], hy[(size_t)j * numelvert + ip], zn};
      const float p4[3] = {h[(size_t)jp * numelvert + ip], hy[(size_t)jp * numelvert + ip], zn};
      if ((p1[0] == 0 && p1[1] == 0) || (p2[0] == 0 && p2[1] == 0) ||
          (p3[0] == 0 && p3[1] == 0) || (p4[0] == 0 && p4[1] == 0))
        return 1 + (long)i * numel + j;
      const size_t o = (size_t)i * W + (j - start);
      x[o] = (float)((p1[0] + p2[0] + p3[0] + p4[0]) / 4.0);
      y[o] = (float)((p1[1] + p2[1] + p3[1] + p4[1]) / 4.0);
      z[o] = (float)((p1[2] + p2[2] + p3[2] + p4[2]) / 4.0);
      float u[3], v[3], w[3], q[3], a1[3], a2[3], b[3];
      for (int c = 0; c < 3; ++c) {
        u[c] = p2[c] - p1[c];
        v[c] = p3[c] - p1[c];
        w[c] = p4[c] - p2[c];
        q[c] = p3[c] - p2[c];
      }
      a1[0] = u[1] * v[2] - v[1] * u[2];
      a1[1] = u[2] * v[0] - v[2] * u[0];
      a1[2] = v[1] * u[0] - u[1] * v[0];
      a2[0] = w[1] * q[2] - q[1] * w[2];
      a2[1] = w[2] * q[0] - q[2] * w[0];
      a2[2] = q[1] * w[0] - w[1] * q[0];
      const float n1 = sqrtf(a1[0] * a1[0] + a1[1] * a1[1] + a1[2] * a1[2]);
      const float n2 = sqrtf(a2[0] * a2[0] + a2[1] * a2[1] + a2[2] * a2[2]);
      area[o] = (float)(0.5 * (n1 + n2));
      for (int c = 0; c < 3; ++c) b[c] = a1[c] + a2[c];
      const float nb = sqrtf(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      nx[o] = b[0] / nb;
      ny[o] = b[1] / nb;
      nz[o] = b[2] / nb;
    }
  }
  return 0;
}

/* util/fields.h:206-234 (average_cpu): phi pass then theta pass, each term
 * divided by the (integer) ratio before it is added, in float. */
void orc_average_grid(const float* in, int nt, int np, int nto, int npo, float* out) {
  const int rp = np / npo, rt = nt / nto;
  float* tmp = (float*)calloc((size_t)nt * npo, sizeof(float));
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < npo; ++j)
      for (int k = 0; k < rp; ++k) tmp[(size_t)i * npo + j] += in[(size_t)i * np + j * rp + k] / (float)rp;
  for (int i = 0; i < nto; ++i)
    for (int j = 0; j < npo; ++j) {
      out[(size_t)i * npo + j] = 0.f;
      for (int k = 0; k < rt; ++k)
        out[(size_t)i * npo + j] += tmp[(size_t)(i * rt + k) * npo + j] / (float)rt;
    }
  free(tmp);
}

/* ------------------------------------------------ 2D cross-section --- */
/* compareBEM2d/rough2d.cu:237-362 (current_fresnel_oblique, eta_in = 1), the
 * sample lengths :477-490, the 2D Green's function :58-76, the sum :520-560 and
 * the projection and normalisation :562-577, for one incidence. */
static void r2_cross(const cplx a[3], const cplx b[3], cplx r[3]) { /* rough2d.cu:172-176 */
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = -(a[0] * b[2] - a[2] * b[0]);
  r[2] = a[0] * b[1] - a[1] * b[0];
}
static cplx r2_dot(const cplx a[3], const cplx b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static cplx r2_rs(double t, cplx nt) { /* rough2d.cu:210-215 with n0 = 1 */
  cplx p0 = 1.0 * cos_internal(1.0, t);
  cplx pt = nt * cos_internal(nt / 1.0, t);
  return (p0 - pt) / (p0 + pt);
}
static cplx r2_rp(double t, cplx nt) { /* rough2d.cu:217-222 with n0 = 1 */
  cplx q0 = cos_internal(1.0, t) / 1.0;
  cplx qt = cos_internal(nt, t) / nt;
  return (q0 - qt) / (q0 + qt);
}

void orc_rough2d(int n, const double* x, const double* y, double theta, double phii,
                 double lambda_m, char mode, double eta_re, double eta_im, int nphi, double Dis,
                 double radius, double* scattering, double* C /* [n][6][2] or NULL */) {
  const double Z0 = orc_z0();
  const cplx eta = eta_re + I * eta_im;
  const double k = 2 * PI / lambda_m, krho = k * cos(theta), omega = k * ORC_C0;
  cplx w[3] = {cos(theta) * cos(phii), cos(theta) * sin(phii), sin(theta)}; /* :264-266 */
  cplx Ed[3], Hd[3];
  const cplx v[3] = {-sin(theta) * cos(phii), -sin(theta) * sin(phii), cos(theta)};
  if (mode == 'M') { /* :268-272, :292 */
    for (int c = 0; c < 3; ++c) Ed[c] = v[c];
    r2_cross(w, Ed, Hd);
  } else { /* :273-277, :296 */
    for (int c = 0; c < 3; ++c) Hd[c] = v[c];
    r2_cross(Hd, w, Ed);
  }
  cplx* cur = (cplx*)calloc((size_t)n * 6, sizeof(cplx));
  double* len = (double*)malloc(sizeof(double) * (size_t)n);
  for (int i = 0; i < n; ++i) {
    double xn = (i < n - 1) ? x[i + 1] : x[0], yn = (i < n - 1) ? y[i + 1] : y[0];
    len[i] = sqrt((x[i] - xn) * (x[i] - xn) + (y[i] - yn) * (y[i] - yn)); /* :486-487 */
    cplx nrm[3] = {yn - y[i], -(xn - x[i]), 0.0};                          /* :303-304 */
    double nn = sqrt(creal(nrm[0] * conj(nrm[0]) + nrm[1] * conj(nrm[1])));
    for (int c = 0; c < 3; ++c) nrm[c] *= 1.0 / nn; /* :305 */
    double phase = creal(k * (w[0] * x[i] + w[1] * y[i]));       /* :286-287 */
    cplx Ei0 = cexp(-I * phase), Hi0 = cexp(-I * phase) / Z0;    /* :288-289 */
    cplx Ei[3], Hi[3], Es[3], Hs[3];
    for (int c = 0; c < 3; ++c) { Ei[c] = Ed[c] * Ei0; Hi[c] = Hd[c] * Hi0; }
    double dotvalue = creal(r2_dot(w, nrm)); /* :307 */
    if (dotvalue <= 0) {
      if (fabs(fabs(dotvalue) - 1) < 1e-15) { /* :311-315 */
        cplx R = r2_rs(0, eta);
        for (int c = 0; c < 3; ++c) { Es[c] = Ei[c] * R; Hs[c] = Hi[c] * R; }
      } else { /* :316-341 */
        cplx q[3], Eperp[3], Epar[3], Rdir[3], Esp[3], Hsp[3], Hip[3], Hsq[3], Esq[3];
        r2_cross(w, nrm, q);
        double qn = sqrt(creal(q[0] * conj(q[0]) + q[1] * conj(q[1]) + q[2] * conj(q[2])));
        for (int c = 0; c < 3; ++c) q[c] *= 1.0 / qn;
        cplx f = r2_dot(Ei, q);
        for (int c = 0; c < 3; ++c) { Eperp[c] = q[c] * f; Epar[c] = Ei[c] - Eperp[c]; }
        double lphi = acos(cabs(r2_dot(w, nrm)));
        cplx R = r2_rs(lphi, eta);
        cplx f2 = 2.0 * r2_dot(w, nrm);
        for (int c = 0; c < 3; ++c) { Rdir[c] = w[c] - nrm[c] * f2; Esp[c] = Eperp[c] * R; }
        r2_cross(Rdir, Esp, Hsp);
        for (int c = 0; c < 3; ++c) Hsp[c] *= 1.0 / Z0;
        R = r2_rp(lphi, eta);
        r2_cross(w, Epar, Hip);
        for (int c = 0; c < 3; ++c) { Hip[c] *= 1.0 / Z0; Hsq[c] = Hip[c] * R; }
        r2_cross(Hsq, Rdir, Esq);
        for (int c = 0; c < 3; ++c) { Esq[c] *= Z0; Es[c] = Esp[c] + Esq[c]; Hs[c] = Hsp[c] + Hsq[c]; }
      }
    } else { /* :344-346 */
      for (int c = 0; c < 3; ++c) { Es[c] = -Ei[c]; Hs[c] = -Hi[c]; }
    }
    cplx Et[3], Ht[3], M[3], J[3];
    for (int c = 0; c < 3; ++c) { Et[c] = Es[c] + Ei[c]; Ht[c] = Hs[c] + Hi[c]; }
    r2_cross(nrm, Et, M); /* :348-350 */
    for (int c = 0; c < 3; ++c) M[c] = -M[c];
    r2_cross(nrm, Ht, J);
    for (int c = 0; c < 3; ++c) { cur[6 * i + c] = J[c]; cur[6 * i + 3 + c] = M[c]; }
  }
  if (C)
    for (int i = 0; i < n; ++i)
      for (int c = 0; c < 6; ++c) {
        C[12 * i + 2 * c] = creal(cur[6 * i + c]);
        C[12 * i + 2 * c + 1] = cimag(cur[6 * i + c]);
      }
  for (int p = 0; p < nphi; ++p) {
    double phi = (double)p / (double)nphi * 2 * PI; /* :521 */
    double dir[3] = {cos(theta) * cos(phi), cos(theta) * sin(phi), sin(theta)};
    cplx E[3] = {0, 0, 0}, H[3] = {0, 0, 0};
    for (int i = 0; i < n; ++i) {
      double d = x[i] * dir[0] + y[i] * dir[1]; /* :37-41 */
      cplx g = csqrt(2.0 * I / (PI * krho * Dis)) * cexp(-I * krho * Dis) * cexp(I * krho * d); /* :69 */
      for (int c = 0; c < 3; ++c) {
        E[c] += cur[6 * i + c] * (-omega * ORC_MU0 / 4.0 * g) * len[i];     /* :89, :114 */
        H[c] += cur[6 * i + 3 + c] * (-omega * ORC_EPS0 / 4.0 * g) * len[i];
      }
    }
    cplx dv = E[0] * dir[0] + E[1] * dir[1] + E[2] * dir[2]; /* :562-569 */
    for (int c = 0; c < 3; ++c) E[c] -= dv * dir[c];
    dv = H[0] * dir[0] + H[1] * dir[1] + H[2] * dir[2];
    for (int c = 0; c < 3; ++c) H[c] -= dv * dir[c];
    cplx Ex = E[0] - (dir[1] * H[2] - dir[2] * H[1]) * Z0; /* :571-573 */
    cplx Ey = E[1] - (-(dir[0] * H[2] - dir[2] * H[0])) * Z0;
    cplx Ez = E[2] - (dir[0] * H[1] - dir[1] * H[0]) * Z0;
    double m2 = creal(Ex * conj(Ex)) + creal(Ey * conj(Ey)) + creal(Ez * conj(Ez));
    scattering[p] = m2 * Dis / (2 * radius); /* :577 */
  }
  free(cur);
  free(len);
}
