// wf_geometry.hip -- fiber heightfield -> surface samples (the
// reference's set_sourcepoint, util/geometry.h:78-178, and its windowless
// form util/geometry.h:3-72).
//
// A block covers a tile of TI rings x TJ arc samples.  The (TJ + 1) x
// (TI + 1) vertices it needs (x and y of arc j + 1 wrapping to 0, ring i + 1
// clamped at the last ring) are loaded once into LDS with consecutive threads
// on consecutive rings (contiguous in the heightfield), then each thread
// forms one quad: lanes run along the arc, so the seven output arrays are
// written contiguously.  The arithmetic follows the reference's float code
// operation by operation (component-wise float products and differences for
// the cross products, left-to-right sums, float square roots and divisions,
// the ring height and the /4 and 0.5 factors through double).  This file is
// compiled without contraction into fused multiply-adds, divisions are
// correctly rounded, and each float square root is taken in double and
// rounded once (the exact root of a float is never close enough to a float
// rounding boundary for the double error to matter), so results are
// bit-identical to a straightforward CPU evaluation.  A vertex with x = y = 0 records the
// smallest i * numel + j of the samples that touch it (the reference reports
// the first such sample in its (i, j) loop and exits).
#include <stdint.h>

#include "wf_common.h"
#include "wf_interp.h"

namespace {

constexpr int TI = 8, TJ = 32;

// correctly rounded float square root
__device__ inline float csqrt(float x) { return (float)sqrt((double)x); }

__global__ __launch_bounds__(256) void heightfield_points_kernel(
    const float* __restrict__ hx, const float* __restrict__ hy, int numel, int numelvert,
    float period, float unitz, int start, int W, float* __restrict__ out, int64_t n,
    int* __restrict__ first_bad) {
  __shared__ float vx[TJ + 1][TI + 1], vy[TJ + 1][TI + 1];
  const int i0 = blockIdx.y * TI, jw0 = blockIdx.x * TJ;  // jw: arc index in the window
  for (int t = threadIdx.x; t < (TJ + 1) * (TI + 1); t += blockDim.x) {
    const int jj = t / (TI + 1), ii = t - jj * (TI + 1);
    // arc offsets beyond the window's last neighbour (W) are not used
    const int jw = (jw0 + jj < W) ? jw0 + jj : W;
    int j = start + jw;
    if (j >= numel) j -= numel;  // the neighbour of arc numel - 1 is arc 0
    int i = i0 + ii;
    if (i > numelvert - 1) i = numelvert - 1;
    const int64_t v = (int64_t)j * numelvert + i;
    vx[jj][ii] = hx[v];
    vy[jj][ii] = hy[v];
  }
  __syncthreads();
  const int tj = threadIdx.x % TJ, ti = threadIdx.x / TJ;
  const int jw = jw0 + tj, i = i0 + ti;
  if (jw >= W || i >= numelvert) return;
  const int j = start + jw;
  // ring i + 1 is clamped at the last ring (same vertices, height z + unitz)
  const int di = (i < numelvert - 1) ? 1 : 0;
  const float z = (float)((double)((float)i * unitz) - (double)period * 0.5);
  const float zn = (float)((double)((float)(i + 1) * unitz) - (double)period * 0.5);
  const float p1x = vx[tj][ti], p1y = vy[tj][ti];
  const float p2x = vx[tj + 1][ti], p2y = vy[tj + 1][ti];
  const float p3x = vx[tj][ti + di], p3y = vy[tj][ti + di];
  const float p4x = vx[tj + 1][ti + di], p4y = vy[tj + 1][ti + di];
  if ((p1x == 0.f && p1y == 0.f) || (p2x == 0.f && p2y == 0.f) || (p3x == 0.f && p3y == 0.f) ||
      (p4x == 0.f && p4y == 0.f))
    atomicMin(first_bad, i * numel + j);
  // centroid
  const float sx = (float)((double)(((p1x + p2x) + p3x) + p4x) / 4.0);
  const float sy = (float)((double)(((p1y + p2y) + p3y) + p4y) / 4.0);
  const float sz = (float)((double)(((z + z) + zn) + zn) / 4.0);
  // a1 = (p2 - p1) x (p3 - p1), a2 = (p4 - p2) x (p3 - p2)
  const float ux = p2x - p1x, uy = p2y - p1y, uz = z - z;
  const float vx1 = p3x - p1x, vy1 = p3y - p1y, vz1 = zn - z;
  const float a1x = uy * vz1 - vy1 * uz;
  const float a1y = uz * vx1 - vz1 * ux;
  const float a1z = vy1 * ux - uy * vx1;
  const float wx = p4x - p2x, wy = p4y - p2y, wz = zn - z;
  const float qx = p3x - p2x, qy = p3y - p2y, qz = zn - z;
  const float a2x = wy * qz - qy * wz;
  const float a2y = wz * qx - qz * wx;
  const float a2z = qy * wx - wy * qx;
  const float n1 = csqrt((a1x * a1x + a1y * a1y) + a1z * a1z);
  const float n2 = csqrt((a2x * a2x + a2y * a2y) + a2z * a2z);
  const float area = (float)(0.5 * (double)(n1 + n2));
  const float bx = a1x + a2x, by = a1y + a2y, bz = a1z + a2z;
  const float nb = csqrt((bx * bx + by * by) + bz * bz);
  const int64_t o = (int64_t)i * W + jw;
  out[o] = sx;
  out[n + o] = sy;
  out[2 * n + o] = sz;
  out[3 * n + o] = bx / nb;
  out[4 * n + o] = by / nb;
  out[5 * n + o] = bz / nb;
  out[6 * n + o] = area;
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

}  // namespace

namespace wf {

hipError_t launch_heightfield_points(const float* height, int numel, int numelvert, float period,
                                     float unitz, int start, int end, float* out, int* first_bad,
                                     hipStream_t st) {
  const int W = end - start + 1;
  const int64_t n = (int64_t)W * numelvert;
  hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, st, first_bad, 0x7fffffff);
  const dim3 grid((unsigned)((W + TJ - 1) / TJ), (unsigned)((numelvert + TI - 1) / TI));
  hipLaunchKernelGGL(heightfield_points_kernel, grid, dim3(256), 0, st, height,
                     height + (int64_t)numel * numelvert, numel, numelvert, period, unitz, start, W,
                     out, n, first_bad);
  return hipGetLastError();
}

}  // namespace wf
