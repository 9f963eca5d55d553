// wf_adjoint.h -- host launchers of the reverse-mode kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wf_common.h"
#include "wf_fft.h"

namespace wf {

struct LeafAdjArgs {
  const float* px;
  const float* py;
  const float* pz;
  const int32_t* leaf_begin;
  const int32_t* leaf_end;
  const uint32_t* leaf_key;
  int nleaves;
  int depth;
  double box_lo[3];
  double leaf_size[3];
  double k0;
  int nt, np;
  int F;
  const float2* g;  // leaf cogradients [leaf][F][nt*np]
  float2* Q;        // [n][F][4]
};

hipError_t launch_epilogue_adj(const double2* F0, const float* gI, int nt, int np, double2* gF,
                               float* Iout, hipStream_t st);
hipError_t launch_theta_adj(const float2* gP, const double2* gProot, int F, int ntc, int npp,
                            int ntp, int nchild, const int32_t* parent_idx,
                            const uint32_t* child_key, const float2* T, double shift,
                            int mirror_base, const FftPlanDev& pp, const FftPlanDev& pc,
                            float2* gG, hipStream_t st);
hipError_t launch_phi_adj(const float2* in, float2* out, int64_t planes, int ntc, int npc, int npp,
                          const FftPlanDev& pp, const FftPlanDev& pc, hipStream_t st);
hipError_t launch_leaf_adj(const LeafAdjArgs& a, hipStream_t st);
hipError_t launch_direct_adj_fp64(const float* pts, int64_t n, const double* centre, double k0,
                                  int nt, int np, const double2* g, double2* Q, hipStream_t st);
hipError_t launch_grad_assemble(const float* pts, int64_t n, const void* Q, bool q_fp64, int F,
                                int f0, const WfIncDev* inc, double* acc, hipStream_t st);
hipError_t launch_grad_scatter(const double* acc, const int32_t* perm, int64_t n, float* gx,
                               float* gn, float* ga, hipStream_t st);

}  // namespace wf
