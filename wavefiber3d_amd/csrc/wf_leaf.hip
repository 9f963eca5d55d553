// wf_leaf.hip -- leaf radiation sum.
//
// Reference: current_times_green_center (util/fields.h:41-50, functor
// :8-38) evaluates C_n exp(+i k s.(x_n - c_leaf(n))) over ALL N points for
// one leaf-grid direction at a time, then one blocking thrust::reduce + one
// fill per (direction, active leaf) (treePO.cu:614-685), ~2 D_leaf L launches
// per solve.  Here the leaf sum is a grouped complex GEMM
//   P[d][n] = exp(+i k s_d.(x_n - c_L))  (D_leaf x n_L, generated on the fly)
//   Y[d][f] = sum_n P[d][n] C[n][f]      (f = 6 components x S solves)
// with one block per (leaf, tile of directions, field block).  The phase uses
// local coordinates (|x - c_L| below the leaf half-diagonal), so FP32 keeps it
// to ~1e-6 rad; the argument is reduced exactly in revolutions.
#include "wf_common.h"
#include "wf_interp.h"

namespace {

constexpr int kLeafThreads = 256;
constexpr int kLeafTile = 64;

__device__ inline void leaf_centre(uint32_t key, int depth, const double lo[3],
                                   const double size[3], double c[3]) {
  int ix = 0, iy = 0, iz = 0;
  for (int l = depth - 1; l >= 0; --l) {
    const uint32_t o = (key >> (3 * l)) & 7u;
    ix = 2 * ix + (int)(o & 1u);
    iy = 2 * iy + (int)((o & 2u) == 0);
    iz = 2 * iz + (int)((o & 4u) == 0);
  }
  c[0] = lo[0] + (ix + 0.5) * size[0];
  c[1] = lo[1] + (iy + 0.5) * size[1];
  c[2] = lo[2] + (iz + 0.5) * size[2];
}

template <int FB>
__global__ __launch_bounds__(kLeafThreads) void leaf_fp32_kernel(wf::LeafArgs a) {
  __shared__ float3 sr[kLeafTile];
  __shared__ float2 sc[kLeafTile][FB];
  const int D = a.nt * a.np;
  const int tiles = (D + kLeafThreads - 1) / kLeafThreads;
  const int nfb = a.F / FB;
  const int64_t bid = blockIdx.x;
  const int tile = (int)(bid % tiles);
  const int fb = (int)((bid / tiles) % nfb) * FB;  // first field of this block within the window
  const int leaf = (int)(bid / ((int64_t)tiles * nfb));
  const int d = tile * kLeafThreads + threadIdx.x;
  const int b = a.leaf_begin[leaf], e = a.leaf_end[leaf];
  double c[3];
  leaf_centre(a.leaf_key[leaf], a.depth, a.box_lo, a.leaf_size, c);
  const double k_rev = a.k0 / WF_TWO_PI;
  float s[3] = {0.f, 0.f, 0.f};
  if (d < D) {
    const int j = d / a.np, kk = d - j * a.np;
    const double th = -WF_PI / 2 + WF_PI * (j + 0.5) / a.nt;
    const double ph = WF_TWO_PI * kk / a.np;
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    s[0] = (float)(ct * cp);
    s[1] = (float)(ct * sp);
    s[2] = (float)st;
  }
  float2 acc[FB];
#pragma unroll
  for (int f = 0; f < FB; ++f) acc[f] = make_float2(0.f, 0.f);
  for (int base = b; base < e; base += kLeafTile) {
    const int cnt = min(kLeafTile, e - base);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += kLeafThreads) {
      const int i = base + t;
      sr[t] = make_float3((float)(((double)a.px[i] - c[0]) * k_rev),
                          (float)(((double)a.py[i] - c[1]) * k_rev),
                          (float)(((double)a.pz[i] - c[2]) * k_rev));
    }
    for (int t = threadIdx.x; t < cnt * FB; t += kLeafThreads) {
      const int p = t / FB, f = t - p * FB;
      sc[p][f] = a.C[(int64_t)(base + p) * a.Ftot + a.f0 + fb + f];
    }
    __syncthreads();
    for (int p = 0; p < cnt; ++p) {
      const float3 r = sr[p];
      float rev = s[0] * r.x + s[1] * r.y + s[2] * r.z;
      rev -= rintf(rev);
      float sn, cs;
      sincosf(rev * 6.28318530717958648f, &sn, &cs);
#pragma unroll
      for (int f = 0; f < FB; ++f) {
        const float2 cc = sc[p][f];
        acc[f].x += cc.x * cs - cc.y * sn;
        acc[f].y += cc.x * sn + cc.y * cs;
      }
    }
  }
  if (d < D) {
#pragma unroll
    for (int f = 0; f < FB; ++f)
      a.out[((int64_t)leaf * a.F + fb + f) * D + d] = acc[f];
  }
}

}  // namespace

namespace wf {

hipError_t launch_leaf(const LeafArgs& a, hipStream_t st) {
  const int D = a.nt * a.np;
  if (a.nleaves <= 0) return hipSuccess;
  if (a.F % 6 != 0 || a.transverse) return hipErrorInvalidValue;
  const unsigned tiles = (unsigned)((D + kLeafThreads - 1) / kLeafThreads);
  const int FB = (a.F % 12 == 0) ? 12 : 6;
  const int64_t nblk = (int64_t)tiles * (a.F / FB) * a.nleaves;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  if (FB == 12)
    hipLaunchKernelGGL(leaf_fp32_kernel<12>, dim3((unsigned)nblk), dim3(kLeafThreads), 0, st, a);
  else
    hipLaunchKernelGGL(leaf_fp32_kernel<6>, dim3((unsigned)nblk), dim3(kLeafThreads), 0, st, a);
  return hipGetLastError();
}

}  // namespace wf
