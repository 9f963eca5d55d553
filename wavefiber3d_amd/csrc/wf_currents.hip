// wf_currents.hip -- tangent-plane Fresnel surface currents (rows A2-A4).
//
// Restates util/tangentplane.h:18-80 (LocalTangentPlane), util/util.h:365-382
// (cos_internal_angle, rs, rp with n0 = 1) and fuses util/fields.h:106-126
// (field_times_area) and :128-150 (multiply_current_factor).  One thread per
// (point, incidence); the math is FP64 (a few hundred flops per point, far
// below the HBM time of the 48 B x S written), so the lit/shadow and
// normal-incidence decisions are the ones the FP64 oracle takes.
#include "wf_common.h"

namespace {

__device__ inline double2 zsqrt(double2 z) {
  // principal branch, same branch cut as C99 csqrt (sign of the imaginary part)
  if (z.x == 0.0 && z.y == 0.0) return make_double2(0.0, z.y);
  double t = sqrt((fabs(z.x) + hypot(z.x, z.y)) * 0.5);
  if (z.x >= 0.0) return make_double2(t, z.y / (2.0 * t));
  return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}

// util.h:365-368 with complex n: sqrt(1 - (sin t / n)^2)
__device__ inline double2 cos_internal(double2 n, double st) {
  double2 s = zdiv(make_double2(st, 0.0), n);
  double2 s2 = zmul(s, s);
  return zsqrt(make_double2(1.0 - s2.x, -s2.y));
}

struct Z3 {
  double2 v[3];
};

__device__ inline Z3 cross_rz(const double a[3], const Z3& b) {
  Z3 r;
  r.v[0] = zsub(zscale(b.v[2], a[1]), zscale(b.v[1], a[2]));
  r.v[1] = zsub(zscale(b.v[0], a[2]), zscale(b.v[2], a[0]));
  r.v[2] = zsub(zscale(b.v[1], a[0]), zscale(b.v[0], a[1]));
  return r;
}
__device__ inline Z3 cross_zr(const Z3& a, const double b[3]) {
  Z3 r;
  r.v[0] = zsub(zscale(a.v[1], b[2]), zscale(a.v[2], b[1]));
  r.v[1] = zsub(zscale(a.v[2], b[0]), zscale(a.v[0], b[2]));
  r.v[2] = zsub(zscale(a.v[0], b[1]), zscale(a.v[1], b[0]));
  return r;
}

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
    OutT* __restrict__ out, int64_t stride_n, int64_t stride_s) {
#pragma clang fp contract(off)
  const int s = blockIdx.y;
  const WfIncDev inc = incs[s];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double p[3] = {px[i], py[i], pz[i]};
    const double nr[3] = {pnx[i], pny[i], pnz[i]};
    const double area = parea[i];
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
      if (fabs(fabs(dotvalue) - 1.0) < 1e-10) {  // :40-49
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
        const double lphi = acos(fabs(dotvalue));   // :58
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
    double2 cur[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      cur[c] = zmul(J.v[c], ef);      // fields.h:123-126, :147-150
      cur[3 + c] = zmul(M.v[c], hf);
    }
    store6<OutT>(out, i * stride_n + (int64_t)s * stride_s, cur);
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
  dim3 grid((unsigned)blocks, (unsigned)n_inc);
  if (fp64)
    hipLaunchKernelGGL(currents_kernel<double2>, grid, dim3(256), 0, st, x, y, z, nx, ny, nz,
                       area, n, inc, (double2*)out, stride_n, stride_s);
  else
    hipLaunchKernelGGL(currents_kernel<float2>, grid, dim3(256), 0, st, x, y, z, nx, ny, nz,
                       area, n, inc, (float2*)out, stride_n, stride_s);
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
