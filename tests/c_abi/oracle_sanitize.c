/*
 * Drives every function of the FP64 C oracle on small inputs, for a build
 * with AddressSanitizer and UndefinedBehaviorSanitizer (gcc, host code only):
 * an out-of-bounds access, use of uninitialised heap memory through a
 * sanitizer-visible path, signed overflow or a misaligned load aborts with a
 * report.  Exit status 0 when every call returns normally and its outputs are
 * finite.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#include "po_oracle.h"

static int bad = 0;
static void finite_or_fail(const char* what, const double* v, long n) {
  for (long i = 0; i < n; ++i)
    if (!isfinite(v[i])) {
      fprintf(stderr, "non-finite output in %s at %ld\n", what, i);
      ++bad;
      return;
    }
}

int main(void) {
  const int n = 500, nt = 6, np = 12, D = nt * np;
  float* p = (float*)malloc(sizeof(float) * 7 * n);
  float *x = p, *y = p + n, *z = p + 2 * n, *nx = p + 3 * n, *ny = p + 4 * n, *nz = p + 5 * n,
        *a = p + 6 * n;
  const double g = M_PI * (3.0 - sqrt(5.0)), r = 1e-6;
  for (int i = 0; i < n; ++i) {
    const double zz = 1.0 - (2.0 * i + 1.0) / n, rr = sqrt(1.0 - zz * zz);
    nx[i] = (float)(rr * cos(g * i));
    ny[i] = (float)(rr * sin(g * i));
    nz[i] = (float)zz;
    x[i] = (float)(r * nx[i]);
    y[i] = (float)(r * ny[i]);
    z[i] = (float)(r * nz[i]);
    a[i] = (float)(4 * M_PI * r * r / n);
  }
  orc_incidence inc[2] = {{M_PI / 2, 0.0, 0.5e-6, 'M', 1.55, 0.1, 0.0},
                          {0.3, 1.0, 0.45e-6, 'E', 1.3, 0.0, 2e-6}};
  double rs[2], rp[2], w[3], E[3], H[3], Ef[2], Hf[2];
  orc_fresnel(0.4, 1.55, 0.1, rs, rp);
  orc_incidence_basis(0.3, 1.0, 'M', w, E, H);
  orc_radiation_factors(0.5e-6, Ef, Hf);
  finite_or_fail("fresnel", rs, 2);
  double* C = (double*)malloc(sizeof(double) * 12 * n * 2);
  orc_currents(n, x, y, z, nx, ny, nz, a, &inc[0], C);
  orc_currents(n, x, y, z, nx, ny, nz, a, &inc[1], C + 12 * n);
  finite_or_fail("currents", C, 24 * n);
  double* dirs = (double*)malloc(sizeof(double) * 3 * D);
  orc_output_dirs(nt, np, dirs);
  double cen[3] = {0, 0, 0};
  double* F = (double*)malloc(sizeof(double) * 12 * D * 2);
  orc_direct_sum(n, x, y, z, cen, 2 * M_PI / 0.5e-6, 2, C, D, dirs, F, 1);
  orc_direct_sum_finite(n, x, y, z, 2 * M_PI / 0.5e-6, 2, C, D, dirs, 1.0, F, 1);
  double* I = (double*)malloc(sizeof(double) * 2 * D);
  double* Eo = (double*)malloc(sizeof(double) * 2 * 6 * D);
  orc_intensity(D, dirs, F, I, Eo);
  if (orc_solve_direct(n, x, y, z, nx, ny, nz, a, inc, 2, nt, np, cen, I, Eo, 1)) ++bad;
  finite_or_fail("solve_direct", I, 2 * D);
  if (orc_solve_direct_dirs(n, x, y, z, nx, ny, nz, a, inc, 2, D, dirs, cen, I, Eo, 1)) ++bad;
  if (orc_solve_finite_dirs(n, x, y, z, nx, ny, nz, a, inc, 2, D, dirs, 1.0, I, Eo, 1)) ++bad;
  finite_or_fail("solve_finite", I, 2 * D);
  const int depth = 3, nodes = (512 * 8 - 1) / 7;
  int64_t* perm = (int64_t*)malloc(sizeof(int64_t) * n);
  int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * nodes);
  int32_t* ne = (int32_t*)malloc(sizeof(int32_t) * nodes);
  float* nbox = (float*)malloc(sizeof(float) * 6 * nodes);
  const float box[6] = {-1.1e-6f, -1.1e-6f, -1.1e-6f, 1.1e-6f, 1.1e-6f, 1.1e-6f};
  if (orc_octree(n, x, y, z, box, depth, perm, nb, ne, nbox)) ++bad;
  if (orc_ref_level_count(3, 36e-6f, 6.0f) <= 0) ++bad;
  double in[2 * 9], out[2 * 20];
  for (int i = 0; i < 18; ++i) in[i] = sin(0.3 * i);
  orc_trig_interp(9, in, 0.1, 20, -0.2, out);
  orc_trig_interp(8, in, 0.0, 5, 0.3, out);
  finite_or_fail("trig_interp", out, 10);
  double* pat = (double*)malloc(sizeof(double) * 2 * 4 * 8);
  double* pout = (double*)malloc(sizeof(double) * 2 * 7 * 12);
  for (int i = 0; i < 64; ++i) pat[i] = cos(0.7 * i);
  orc_sphere_interp(4, 8, -M_PI / 2 + M_PI / 8, pat, 7, 12, -M_PI / 2 + M_PI / 14, pout);
  finite_or_fail("sphere_interp", pout, 2 * 7 * 12);
  const int numel = 16, numelvert = 5;
  float* hf = (float*)malloc(sizeof(float) * 2 * numel * numelvert);
  for (int j = 0; j < numel; ++j)
    for (int i = 0; i < numelvert; ++i) {
      hf[j * numelvert + i] = (float)(2e-6 * cos(2 * M_PI * j / numel));
      hf[numel * numelvert + j * numelvert + i] = (float)(1.5e-6 * sin(2 * M_PI * j / numel));
    }
  const int hn = (numel - 3) * numelvert;
  float* hs = (float*)malloc(sizeof(float) * 7 * hn);
  if (orc_heightfield_points(hf, numel, numelvert, 1e-6f, 0.2e-6f, 1, numel - 3, hs, hs + hn,
                             hs + 2 * hn, hs + 3 * hn, hs + 4 * hn, hs + 5 * hn, hs + 6 * hn) != 0)
    ++bad;
  float grid[12 * 20], avg[3 * 4];
  for (int i = 0; i < 240; ++i) grid[i] = (float)i;
  orc_average_grid(grid, 12, 20, 3, 4, avg);
  const int m = 64;
  double* cx = (double*)malloc(sizeof(double) * m);
  double* cy = (double*)malloc(sizeof(double) * m);
  for (int i = 0; i < m; ++i) {
    cx[i] = 5e-6 * cos(2 * M_PI * i / m);
    cy[i] = 4e-6 * sin(2 * M_PI * i / m);
  }
  double* sc = (double*)malloc(sizeof(double) * 90);
  double* cc = (double*)malloc(sizeof(double) * 12 * m);
  orc_rough2d(m, cx, cy, 0.0, M_PI, 0.5e-6, 'M', 1.55, 0.1, 90, 1.0, 10e-6, sc, cc);
  orc_rough2d(m, cx, cy, 0.2, M_PI, 0.6e-6, 'E', 1.55, 0.1, 90, 1.0, 10e-6, sc, NULL);
  finite_or_fail("rough2d", sc, 90);
  free(p); free(C); free(dirs); free(F); free(I); free(Eo); free(perm); free(nb); free(ne);
  free(nbox); free(pat); free(pout); free(hf); free(hs); free(cx); free(cy); free(sc); free(cc);
  if (bad) return 1;
  printf("oracle under ASan/UBSan: ok\n");
  return 0;
}
