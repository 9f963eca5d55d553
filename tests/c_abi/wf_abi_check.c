/*
 * A C caller of the C-ABI: includes include/wf_po.h and links libwfpo.so,
 * as a reference driver would (the compareMie/sphere.cu body of the
 * integration notes).
 *
 *   wf_abi_check errors   argument validation and error reporting only
 *                         (no device work; runs without a GPU)
 *   wf_abi_check solve    one tree solve of a 3000-point Fibonacci sphere
 *                         (depth 2, TM and TE, 24 x 48) against the FP64
 *                         direct sum: relative L2 of intensity <= 1e-3
 *
 * Exit status 0 on success; every failure prints the check that failed.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#include "wf_po.h"

static int failures = 0;

#define CHECK(cond, what)                                            \
  do {                                                               \
    if (!(cond)) {                                                   \
      fprintf(stderr, "FAILED: %s (last error: %s)\n", what, wf_last_error()); \
      ++failures;                                                    \
    }                                                                \
  } while (0)

static int errors_mode(void) {
  wf_opts o;
  wf_handle* h = NULL;
  float dummy[8] = {0};
  float out[8];
  wf_grid g = {4, 8};
  wf_incidence inc = {0.0, 0.0, 0.5e-6, 'M', 1.55, 0.0, 0.0};
  CHECK(wf_abi_version() == WF_ABI_VERSION, "abi version");
  wf_default_opts(&o);
  CHECK(o.struct_size == (int32_t)sizeof(wf_opts), "default opts struct_size");
  CHECK(o.beta > 0 && o.box_enlarge > 0, "default opts values");
  CHECK(wf_create(NULL, &o) == WF_ERR_INVALID_ARG, "wf_create(NULL)");
  CHECK(strlen(wf_last_error()) > 0, "error message set");
  o.struct_size = 8;
  CHECK(wf_create(&h, &o) == WF_ERR_INVALID_ARG && h == NULL, "opts struct_size mismatch");
  CHECK(strstr(wf_last_error(), "struct_size") != NULL, "struct_size message");
  CHECK(wf_solve(NULL, &inc, 1, g, out, NULL, WF_MEM_HOST, NULL) == WF_ERR_INVALID_ARG,
        "wf_solve(NULL handle)");
  CHECK(wf_set_geometry(NULL, NULL, 1, NULL) == WF_ERR_INVALID_ARG, "wf_set_geometry(NULL)");
  CHECK(wf_get_stats(NULL, NULL) == WF_ERR_INVALID_ARG, "wf_get_stats(NULL)");
  CHECK(wf_average_grid(NULL, 1, 4, 8, 2, 4, out, WF_MEM_HOST, NULL) == WF_ERR_INVALID_ARG,
        "wf_average_grid(NULL input)");
  CHECK(wf_average_grid(dummy, 1, 2, 4, 4, 8, out, WF_MEM_HOST, NULL) == WF_ERR_INVALID_ARG,
        "wf_average_grid output finer than input");
  CHECK(wf_heightfield_points(NULL, 4, 4, 1e-6f, 0.25e-6f, 0, 3, out, WF_MEM_HOST, NULL) ==
            WF_ERR_INVALID_ARG,
        "wf_heightfield_points(NULL)");
  CHECK(wf_read_heightfield("/nonexistent/fiber0.binary", 4, 4, out) != WF_OK,
        "wf_read_heightfield(missing file)");
  CHECK(wf_batch_solve(NULL, &inc, 1, g, out, WF_MEM_HOST, NULL) == WF_ERR_INVALID_ARG,
        "wf_batch_solve(NULL)");
  CHECK(wf_rough2d_solve(NULL, NULL, 0, NULL, 0, 0, 1.0, 1e-5, NULL, WF_MEM_HOST, NULL) ==
            WF_ERR_INVALID_ARG,
        "wf_rough2d_solve(NULL)");
  wf_destroy(NULL); /* a no-op */
  if (!failures) printf("errors: ok\n");
  return failures ? 1 : 0;
}

static int solve_mode(void) {
  const int n = 3000, nt = 24, np = 48;
  const double r = 1.5e-6, golden = M_PI * (3.0 - sqrt(5.0));
  float* buf = (float*)malloc(sizeof(float) * 7 * n);
  float *x = buf, *y = buf + n, *z = buf + 2 * n, *nx = buf + 3 * n, *ny = buf + 4 * n,
        *nz = buf + 5 * n, *area = buf + 6 * n;
  for (int i = 0; i < n; ++i) {
    const double zz = 1.0 - (2.0 * i + 1.0) / n, rr = sqrt(1.0 - zz * zz), p = golden * i;
    nx[i] = (float)(rr * cos(p));
    ny[i] = (float)(rr * sin(p));
    nz[i] = (float)zz;
    x[i] = (float)(r * rr * cos(p));
    y[i] = (float)(r * rr * sin(p));
    z[i] = (float)(r * zz);
    area[i] = (float)(4.0 * M_PI * r * r / n);
  }
  wf_surface s = {x, y, z, nx, ny, nz, area, n, WF_MEM_HOST};
  wf_incidence inc[2] = {{M_PI / 2, 0.0, 0.5e-6, 'M', 1.55, 0.1, 0.0},
                         {0.3, 1.0, 0.5e-6, 'E', 1.55, 0.1, 0.0}};
  wf_grid g = {nt, np};
  float* I = (float*)malloc(sizeof(float) * 2 * nt * np);
  double* Id = (double*)malloc(sizeof(double) * 2 * nt * np);
  wf_opts o;
  wf_default_opts(&o);
  wf_handle* h = NULL;
  CHECK(wf_create(&h, &o) == WF_OK, "wf_create");
  if (h) {
    CHECK(wf_po_solve(h, &s, inc, 2, g, 2, I, NULL, NULL) == WF_OK, "wf_po_solve");
    CHECK(wf_po_solve_fp64_direct(h, &s, inc, 2, g, Id, NULL, NULL) == WF_OK,
          "wf_po_solve_fp64_direct");
    wf_stats st;
    CHECK(wf_get_stats(h, &st) == WF_OK, "wf_get_stats");
    for (int k = 0; k < 2 && !failures; ++k) {
      double num = 0, den = 0;
      for (int i = 0; i < nt * np; ++i) {
        const double d = I[k * nt * np + i] - Id[k * nt * np + i];
        num += d * d;
        den += Id[k * nt * np + i] * Id[k * nt * np + i];
      }
      const double e = sqrt(num / den);
      printf("solve %d: rel-L2 vs FP64 direct %.3e\n", k, e);
      CHECK(e <= 1e-3, "tree vs FP64 direct");
    }
    wf_destroy(h);
  }
  free(buf);
  free(I);
  free(Id);
  if (!failures) printf("solve: ok\n");
  return failures ? 1 : 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && strcmp(argv[1], "solve") == 0) return solve_mode();
  return errors_mode();
}
