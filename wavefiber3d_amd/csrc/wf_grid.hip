// wf_grid.hip -- block average of output grids (the reference's
// average_cpu, util/fields.h:206-234).
//
// Output sample (i, j) of grid g averages rows i*rt .. i*rt + rt - 1 and
// columns j*rp .. j*rp + rp - 1 with rt = n_theta / n_theta_out and rp =
// n_phi / n_phi_out (integer quotients, as there).  The reference first
// averages along phi for every input row, accumulating x / rp term by term in
// float, then along theta, accumulating row / rt term by term; one thread per
// output sample repeats exactly that sequence (divisions correctly rounded,
// no reassociation), so the result is bit-identical.  Lanes take consecutive
// j, so each of the rp-wide column groups is read contiguously.
#include <stdint.h>

#include "wf_common.h"
#include "wf_interp.h"

namespace {

__global__ __launch_bounds__(256) void average_grid_kernel(const float* __restrict__ in, int ngrids,
                                                           int nt, int np, int nto, int npo,
                                                           int rt, int rp, float* __restrict__ out) {
  const int64_t total = (int64_t)ngrids * nto * npo;
  const float frp = (float)rp, frt = (float)rt;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int j = (int)(t % npo);
    const int64_t r = t / npo;
    const int i = (int)(r % nto);
    const int64_t g = r / nto;
    const float* src = in + (g * nt + (int64_t)i * rt) * np + (int64_t)j * rp;
    float acc = 0.f;
    for (int k = 0; k < rt; ++k) {
      float s = 0.f;
      for (int m = 0; m < rp; ++m) s += src[(int64_t)k * np + m] / frp;
      acc += s / frt;
    }
    out[t] = acc;
  }
}

}  // namespace

namespace wf {

hipError_t launch_average_grid(const float* in, int ngrids, int nt, int np, int nto, int npo,
                               float* out, hipStream_t st) {
  const int64_t total = (int64_t)ngrids * nto * npo;
  if (total <= 0) return hipSuccess;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(average_grid_kernel, dim3((unsigned)blocks), dim3(256), 0, st, in, ngrids, nt,
                     np, nto, npo, nt / nto, np / npo, out);
  return hipGetLastError();
}

}  // namespace wf
