// wf_tree64.h -- launchers of the FP64 tree (wf_tree64.hip): leaf sums, FFT
// level transitions and translation tables in FP64, and their adjoints.
// Internal header; layouts as in the FP32 FFT path (wf_interp.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wf {

// Stockham plan: radices (4 bits per stage, 2/3/4/5/7/11/13) and the FP64
// twiddle table tw[m] = exp(-2 pi i m / n), m < n (device memory).
struct Fft64Plan {
  int n;
  int nstages;
  uint64_t radix_packed;
  const double2* tw;
};

struct Leaf64Args {
  const float* px;  // sorted positions
  const float* py;
  const float* pz;
  const double2* C;  // FP64 currents [n][F] (sorted order)
  int F;             // fields (a multiple of 6)
  const int32_t* leaf_begin;
  const int32_t* leaf_end;
  const uint32_t* leaf_key;
  int nleaves;
  int depth;
  double box_lo[3];
  double leaf_size[3];
  double k0;
  int nt, np;     // leaf grid
  double2* out;   // [leaf][F][nt*np]
};

struct LeafAdj64Args {
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
  const double2* g;  // leaf cogradients [leaf][F][nt*np]
  double2* Q;        // [n][F][4]: sum_d conj(P) {1, s_x, s_y, s_z} g
};

hipError_t launch_leaf64(const Leaf64Args& a, hipStream_t st);
// child rows [plane][ntc][npc] -> [plane][ntc][npp]
hipError_t launch_phi64(const double2* in, double2* out, int64_t planes, int ntc, int npc,
                        int npp, const Fft64Plan& pc, const Fft64Plan& pp, hipStream_t st);
// G [child][F][ntc][npp] -> out [parent][F][ntp][npp] (the root: one parent),
// great-circle theta interpolation, translation by T [8][ntp][npp], child sum
hipError_t launch_theta64(const double2* G, int F, int ntc, int npp, int ntp, int nparents,
                          const int32_t* cb, const int32_t* ce, const uint32_t* child_key,
                          const double2* T, double shift, int mirror_base, const Fft64Plan& pc,
                          const Fft64Plan& pp, double2* out, hipStream_t st);
hipError_t launch_trans64(double2* T, int ntp, int npp, double theta0, const double h[3], double k0,
                          hipStream_t st);
hipError_t launch_theta_adj64(const double2* gP, int F, int ntc, int npp, int ntp, int nchild,
                              const int32_t* parent_idx, const uint32_t* child_key,
                              const double2* T, double shift, int mirror_base,
                              const Fft64Plan& pp, const Fft64Plan& pc, double2* gG,
                              hipStream_t st);
hipError_t launch_phi_adj64(const double2* in, double2* out, int64_t planes, int ntc, int npc,
                            int npp, const Fft64Plan& pp, const Fft64Plan& pc, hipStream_t st);
hipError_t launch_leaf_adj64(const LeafAdj64Args& a, hipStream_t st);

}  // namespace wf
