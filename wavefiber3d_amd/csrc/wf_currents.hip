// wf_currents.hip -- tangent-plane Fresnel surface currents.
//
// Restates util/tangentplane.h:18-80 (LocalTangentPlane), util/util.h:365-382
// (cos_internal_angle, rs, rp with n0 = 1) and fuses util/fields.h:106-126
// (field_times_area) and :128-150 (multiply_current_factor).  One thread per
// point runs every incidence (incidences split over a few block rows only
// when the points alone fill fewer than 1024 blocks), so with the
// point-major layout its stores fill whole cache lines (one thread per
// (point, incidence) wrote 48 B pieces of lines shared with other blocks:
// 1.7x the time on the 663k-point fiber); the
// math is FP64 (a few hundred flops per point and incidence), so the
// lit/shadow and normal-incidence decisions are the ones the FP64 oracle
// takes.
#include "wf_common.h"
#include "wf_currents.h"

namespace {

template <typename OutT>
__device__ inline void store6(OutT* o, int64_t base, const double2 c[6]);

template <>
__device__ inline void store6<float2>(float2* o, int64_t base, const double2 c[6]) {
#pragma unroll
  for (int q = 0; q < 6; ++q) o[base + q] = make_float2((float)c[q].x, (float)c[q].y);
}
template <>
__device__ inline void store6<double2>(double2* o, int64_t base, const double2 c[6]) {
#pragma unroll
  for (int q = 0; q < 6; ++q) o[base + q] = c[q];
}

template <typename OutT>
__global__ __launch_bounds__(256) void currents_kernel(
    const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ pnx, const float* __restrict__ pny, const float* __restrict__ pnz,
    const float* __restrict__ parea, int64_t n, const WfIncDev* __restrict__ incs,
    OutT* __restrict__ out, int64_t stride_n, int64_t stride_s, int n_inc) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double p[3] = {px[i], py[i], pz[i]};
    const double nr[3] = {pnx[i], pny[i], pnz[i]};
    const double ar = parea[i];
    // the point's incidences blockIdx.y, + gridDim.y, ... from one thread
    // (all of them unless the point grid is small): with the point-major
    // layout (stride_s = 6) its stores fill whole cache lines
    for (int s = blockIdx.y; s < n_inc; s += gridDim.y) {
      double2 cur[6];
      wfdev::current_point(p, nr, ar, incs[s], cur);
      store6<OutT>(out, i * stride_n + (int64_t)s * stride_s, cur);
    }
  }
}

// Input validation and bounding box: counts non-finite values, normals far
// from unit length and negative areas; bbox via float atomics on ordered ints.
__device__ inline int ord_of(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}

__global__ __launch_bounds__(256) void validate_kernel(
    const float* __restrict__ x, const float* __restrict__ y, const float* __restrict__ z,
    const float* __restrict__ nx, const float* __restrict__ ny, const float* __restrict__ nz,
    const float* __restrict__ a, int64_t n, int* __restrict__ bbox_ord, int* __restrict__ bad) {
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  int nbad = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float p[3] = {x[i], y[i], z[i]};
    const float q[3] = {nx[i], ny[i], nz[i]};
    const float ar = a[i];
    bool ok = isfinite(p[0]) && isfinite(p[1]) && isfinite(p[2]) && isfinite(q[0]) &&
              isfinite(q[1]) && isfinite(q[2]) && isfinite(ar) && ar >= 0.0f;
    const float nn = q[0] * q[0] + q[1] * q[1] + q[2] * q[2];
    ok = ok && fabsf(nn - 1.0f) < 1e-2f;
    if (!ok) {
      ++nbad;
      continue;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int o = ord_of(p[c]);
      lo[c] = min(lo[c], o);
      hi[c] = max(hi[c], o);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    for (int off = 32; off > 0; off >>= 1) {
      lo[c] = min(lo[c], __shfl_xor(lo[c], off));
      hi[c] = max(hi[c], __shfl_xor(hi[c], off));
    }
  }
  for (int off = 32; off > 0; off >>= 1) nbad += __shfl_xor(nbad, off);
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      atomicMin(&bbox_ord[c], lo[c]);
      atomicMax(&bbox_ord[3 + c], hi[c]);
    }
    if (nbad) atomicAdd(bad, nbad);
  }
}

__global__ void validate_init(int* bbox_ord, int* bad) {
  if (threadIdx.x < 3) {
    bbox_ord[threadIdx.x] = INT_MAX;
    bbox_ord[3 + threadIdx.x] = INT_MIN;
  }
  if (threadIdx.x == 0) *bad = 0;
}

__global__ void validate_fini(const int* bbox_ord, float* bbox) {
  if (threadIdx.x < 6) {
    int o = bbox_ord[threadIdx.x];
    int i = o >= 0 ? o : o ^ 0x7fffffff;
    bbox[threadIdx.x] = __int_as_float(i);
  }
}

}  // namespace

namespace wf {

hipError_t launch_currents(const float* x, const float* y, const float* z, const float* nx,
                           const float* ny, const float* nz, const float* area, int64_t n,
                           const WfIncDev* inc, int n_inc, int fp64, void* out, int64_t stride_n,
                           int64_t stride_s, hipStream_t st) {
  if (n <= 0 || n_inc <= 0) return hipSuccess;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 65535) blocks = 65535;
  // incidence groups only where the points alone give fewer than 1024 blocks
  const int64_t groups = std::min<int64_t>(n_inc, std::max<int64_t>(1, 1024 / blocks));
  dim3 grid((unsigned)blocks, (unsigned)groups);
  if (fp64)
    hipLaunchKernelGGL(currents_kernel<double2>, grid, dim3(256), 0, st, x, y, z, nx, ny, nz,
                       area, n, inc, (double2*)out, stride_n, stride_s, n_inc);
  else
    hipLaunchKernelGGL(currents_kernel<float2>, grid, dim3(256), 0, st, x, y, z, nx, ny, nz,
                       area, n, inc, (float2*)out, stride_n, stride_s, n_inc);
  return hipGetLastError();
}

hipError_t launch_validate(const float* x, const float* y, const float* z, const float* nx,
                           const float* ny, const float* nz, const float* area, int64_t n,
                           float* bbox, int* bad, hipStream_t st) {
  // bbox doubles as scratch for the ordered-int reduction (6 ints then 6 floats)
  int* ord = reinterpret_cast<int*>(bbox) + 8;
  hipLaunchKernelGGL(validate_init, dim3(1), dim3(64), 0, st, ord, bad);
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(validate_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x, y, z, nx, ny,
                     nz, area, n, ord, bad);
  hipLaunchKernelGGL(validate_fini, dim3(1), dim3(64), 0, st, ord, bbox);
  return hipGetLastError();
}

}  // namespace wf
