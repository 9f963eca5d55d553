// wf_interp.h -- host launchers of the aggregation kernels and leaf stage.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "wf_fft.h"
#include "wf_fft2.h"

namespace wf {

hipError_t launch_phi_pass(const float2* in, float2* out, int64_t planes, int ntc, int npc, int npp,
                           const FftPlanDev& pc, const FftPlanDev& pp, hipStream_t st);

// cb/ce: per-parent child range in the child list (nullable: parent p <- child p).
// child_key: level-local node id of each child (octant = key & 7), nullable.
// T: translation table [8][ntp][npp] (nullable: no translation).
// Exactly one of out / out_root is non-null.
hipError_t launch_theta_pass(const float2* G, int F, int ntc, int npp, int ntp, int nparents,
                             const int32_t* cb, const int32_t* ce, const uint32_t* child_key,
                             const float2* T, double shift, int mirror_base,
                             const FftPlanDev& pc, const FftPlanDev& pp, float2* out,
                             double2* out_root, hipStream_t st);

// Fused level transition (phi + theta + translation + child sum) for grids
// whose phi-interpolated child and batch buffers fit in LDS.
bool fused_fits(int ntc, int npc, int ntp, int npp);
hipError_t launch_agg_fused(const float2* child, float2* parent, int F, int ntc, int npc, int ntp,
                            int npp, int nparents, const int32_t* cb, const int32_t* ce,
                            const uint32_t* child_key, const float2* T, double shift,
                            int mirror_base, const FftPlanDev& fpc, const FftPlanDev& fpp,
                            const FftPlanDev& ftc, const FftPlanDev& ftp, hipStream_t st);

hipError_t launch_trans_table(float2* T, int ntp, int npp, double theta0, const double h[3],
                              double k0, hipStream_t st);

// Dense MFMA level transition (wf_agg_mfma.hip).  Interpolation matrices are
// real [rows_pad][kpad] (row q = output sample, column j = input sample),
// stored as split BF16 bit patterns (hi, lo).
void build_interp_matrix(int nc, int np, double shift, std::vector<uint16_t>& hi,
                         std::vector<uint16_t>& lo, int& rows_pad, int& kpad);
int interp_rows_pad(int rows);
int interp_kpad(int k);

// phi pass: X [rows = nodes*F*ntc][npc] -> GT, pre-split great-circle rows:
// for each plane cf = (node, field) and column pair k < npp/2 one row of
// gc_row_len(ntc) = 2 gc_half_len(ntc) samples, j < gc_half_len: column k at
// theta j, then column k + npp/2 at theta j - gc_half_len (both ascending),
// zero where theta >= ntc.  Each group of 8 samples is stored as 4 planes of
// 8 BF16 (re hi, re lo, im hi, im lo; hi + lo = the FP32 value), i.e.
// 64 B per group and 8 B per sample.  npc <= kPhiMfmaMaxK, npp even.
// row_tile: rows per staged tile for rows of more than 160 samples (0 or 32:
// 32; 16: half the LDS per block)
hipError_t launch_phi_mfma(const float2* X, float2* GT, int64_t rows, int ntc, int npc, int npp,
                           const uint16_t* Mhi, const uint16_t* Mlo, int kpad, hipStream_t st,
                           int row_tile = 0);
inline int gc_half_len(int ntc) { return ((ntc + 7) / 8) * 8; }
inline int gc_row_len(int ntc) { return 2 * gc_half_len(ntc); }
// samples (8 B each) of the phi-pass output for `planes` = nodes*F
inline int64_t gc_elems(int64_t planes, int ntc, int npp) {
  return planes * (int64_t)(npp / 2) * gc_row_len(ntc);
}
// the theta matrix in the same column order: [rows_pad >= 2 ntp][kpad >=
// gc_row_len(ntc)], column j < gc_half_len -> great-circle sample j, column
// gc_half_len + t -> sample 2 ntc - 1 - t, zero for padding
// flip: for transverse components (see launch_project_transverse): the
// columns of the second half of the great circle and the rows e >= n2p/2 are
// negated (the transverse unit vectors change sign across the pole)
void build_theta_matrix_gc(int ntc, int n2p, double shift, std::vector<uint16_t>& hi,
                           std::vector<uint16_t>& lo, int& rows_pad, int& kpad, bool flip = false);
// largest npc of the phi MFMA pass and largest gc_row_len(ntc) of the theta
// MFMA pass (their matrix fragments stay in registers for all K steps)
constexpr int kPhiMfmaMaxK = 288;
constexpr int kThetaMfmaMaxK = 512;

struct ThetaMfmaArgs {
  const float2* GT;  // phi_mfma output (pre-split great-circle rows, see above)
  int F, ntc, npp, ntp, nparents;
  const int32_t* cb;  // per-parent child range (null: parent p <- child p)
  const int32_t* ce;
  const uint32_t* child_key;  // octant = key & 7 (nullable)
  const float2* Text;         // [8][2 ntp + 1][npp/2] (null: no translation)
  int text_per_child = 0;     // 1: Text is [child][2 ntp + 1][npp/2] (launch_trans_child)
  const uint16_t* Mhi;        // build_theta_matrix_gc matrix
  const uint16_t* Mlo;
  int kpad;
  int mirror;  // 2 ntp - 1 (half-sample grid) or 2 ntp (sample at the pole)
  float2* out;       // [parent][F][ntp][npp], or
  double2* out_root; // FP64 accumulation target of the root level
  int tile = 0;      // 6-9 K steps: 0/1 = 128 x 32, 2 = 128 x 16, 3 = 64 x 16
  int flip = 0;      // transverse components: the pole's second sample is negated
                     // (the matrix carries the other signs, build_theta_matrix_gc)
};
hipError_t launch_theta_mfma(const ThetaMfmaArgs& a, hipStream_t st);

// One internal level transition in a single kernel (wf_agg_fused.hip): phi
// and theta interpolation with the split-BF16 matrices of the two-pass MFMA
// transition, translation and child sum, the phi output kept in LDS.
// Shapes: npc, npp even, ceil(npc/32) <= 4, ceil(gc_row_len(ntc)/32) <= 5,
// 2 ntp <= 256; internal levels only (mirror = 2 ntp - 1).
struct AggFusedArgs {
  const float2* X;  // child patterns [child][F][ntc][npc]
  float2* out;      // parent patterns [parent][F][ntp][npp]
  int F, ntc, npc, ntp, npp, nparents;
  const int32_t* cb;  // per-parent child range
  const int32_t* ce;
  const uint32_t* child_key;  // octant = key & 7
  const float2* Text;         // [8][2 ntp + 1][npp/2] (launch_trans_ext)
  const uint16_t* Mphi_hi;    // build_interp_matrix(npc, npp, 0)
  const uint16_t* Mphi_lo;
  int kpad_phi, mrows_phi;
  const uint16_t* Mth_hi;  // build_theta_matrix_gc(ntc, 2 ntp, shift)
  const uint16_t* Mth_lo;
  int kpad_th, mrows_th;
  int mirror;
  int version = 1;  // 1: phi operands from global memory; 2: LDS-staged, one child ahead
};
bool agg_fused_fits(int ntc, int npc, int ntp, int npp);
hipError_t launch_agg_fused_mfma(const AggFusedArgs& a, hipStream_t st);

// kmajor = 1: the same tables stored [8][npp/2][2 ntp + 1] (great circle k
// contiguous), the layout of the wave-private LDS FFT theta kernel
hipError_t launch_trans_ext(float2* Text, int ntp, int npp, int mirror, double theta0,
                            const double h[3], double k0, hipStream_t st, int kmajor = 0);
// heightfield -> 7 SoA arrays (wf_geometry.hip); *first_bad receives the
// smallest i * numel + j of a sample with a zero vertex (INT_MAX if none)
hipError_t launch_heightfield_points(const float* height, int numel, int numelvert, float period,
                                     float unitz, int start, int end, float* out, int* first_bad,
                                     hipStream_t st);
// block average of [ngrids][nt][np] onto [ngrids][nto][npo] (wf_grid.hip)
hipError_t launch_average_grid(const float* in, int ngrids, int nt, int np, int nto, int npo,
                               float* out, hipStream_t st);
// one incidence of the 2D cross-section solve (wf_rough2d.hip), FP64, host-made
struct Rough2dIncDev {
  double w[3], Ed[3], Hd[3];  // propagation direction, E and H directions
  double k, krho, ct, st;     // k0, k0 cos t, cos t, sin t (t: elevation)
  double z0, eta_re, eta_im;
  double fe, fh;              // -w mu0/4 and -w eps0/4
};
// 2D cross-section PO: currents C[n_inc][n][6] and partial sums
// part[n_inc][P][6][nphi] (scratch, P = rough2d_split), intensity
// I[n_inc][nphi]; *bad receives the smallest index of a zero-length or
// non-finite segment
int rough2d_split(int n, int n_inc, int nphi);
hipError_t launch_rough2d(const double* x, const double* y, int n, const Rough2dIncDev* incs,
                          int n_inc, int nphi, double dist, double radius, double2* C,
                          double2* part, int P, double* I, int* bad, hipStream_t st);
// per-child extended table for children at offsets (cx, cy, cz)[c] from the
// parent centre; offsets of level-r nodes from the root centre from their keys
// (hx/hy/hz[l]: the child-centre offset magnitudes of the level-l transition)
hipError_t launch_trans_child(float2* T, int ntp, int npp, int mirror, double theta0,
                              const double* cx, const double* cy, const double* cz, int nch,
                              double k0, hipStream_t st, int kmajor = 0);
hipError_t launch_child_offsets(const uint32_t* keys, int n, int r, const double* hx,
                                const double* hy, const double* hz, double* cx, double* cy,
                                double* cz, hipStream_t st);

// FFT level transition for large grids (wf_agg_fft.hip).  fac tables hold
// w_f exp(i f shift) / nc for f = -(nc/2) .. band top (nc + 1 entries).
hipError_t launch_phi_fft(const float2* X, float2* GT, int64_t lines_cf, int ntc, int npc, int npp,
                          const Fft2Plan& pc, const Fft2Plan& pp, const float2* fac,
                          hipStream_t st);
struct ThetaFftArgs {
  const float2* GT;  // [child][F][npp][ntc]
  int F, ntc, npp, ntp, nparents;
  const int32_t* cb;
  const int32_t* ce;
  const uint32_t* child_key;
  const float2* Text;  // [8][2 ntp + 1][npp/2] (nullable)
  int text_per_child = 0;  // 1: Text is [child][2 ntp + 1][npp/2]
  int mirror;
  Fft2Plan pc;         // length 2 ntc
  Fft2Plan pp;         // length 2 ntp
  const float2* fac;   // resample factors, 2 ntc + 1
  float2* out;
  double2* out_root;
  int flip = 0;        // transverse components: the second half of every great
                       // circle is negated on input and on output
  int text_kmajor = 0; // Text stored [o][npp/2][2 ntp + 1] (required by the
                       // wave-private kernel, theta_fft_wave)
};
hipError_t launch_theta_fft(const ThetaFftArgs& a, hipStream_t st);

// theta stride of the FFT phi output G [child][F][npp][fft_g_stride(ntc)]:
// ntc rounded up to 16 samples, so 16-row phi blocks write whole 128-byte
// lines and every column starts line-aligned (the padding is never read)
__host__ __device__ inline int fft_g_stride(int ntc) { return (ntc + 15) & ~15; }

// The LDS FFT kernels keep lines of up to 1024 samples wave-private (one wave
// per line, no workgroup barriers between stages); longer lines run the
// block-wide kernels.  The theta choice fixes the translation-table layout.
inline bool phi_fft_wave(int npc, int npp) {
  return (npc > npp ? npc : npp) <= 1024 && npp >= 4;
}
inline bool theta_fft_wave(int ntc, int ntp) {
  return (ntc > ntp ? 2 * ntc : 2 * ntp) <= 1024 && ntp >= 2;
}

// Transverse components of the far field from the six radiated fields of each
// solve: F [node][6 S][nt][np] (E x, y, z, H x, y, z; level grid theta_j =
// theta0 + pi j / nt, phi_k = 2 pi k / np) -> P [node][2 S][nt][np] with
//   P[2s]   = t.E - z0 p.H,   P[2s+1] = p.E + z0 t.H,
// t = (-sin th cos ph, -sin th sin ph, cos th), p = (-sin ph, cos ph, 0): the
// components along t and p of E'' = E_perp - z0 s x H_perp, the field whose
// squared norm is the intensity.  Each is a band-limited function of
// (theta, phi) of one degree more than E and H, and changes sign with t and p
// across the pole, which the theta passes' flip handles.
hipError_t launch_project_transverse(const float2* F, float2* P, int64_t nodes, int S, int nt,
                                     int np, double theta0, double z0, hipStream_t st);

// Leaf radiation sum: pattern[leaf][F][nt][np] of
// sum_{n in leaf} C[n][f] exp(+i k0 s.(x_n - c_leaf)) on the leaf grid
// theta_j = -pi/2 + pi (j + 1/2)/nt, phi_k = 2 pi k/np.
struct LeafArgs {
  const float* px;  // sorted positions (x, y, z consecutive arrays of n)
  const float* py;
  const float* pz;
  int64_t n;
  const float2* C;  // currents [n][Ftot] (sorted order)
  int Ftot;         // fields per point in C
  int f0, F;        // field window handled (pattern holds F fields)
  const int32_t* leaf_begin;
  const int32_t* leaf_end;
  const uint32_t* leaf_key;
  int nleaves;
  int depth;
  double box_lo[3];
  double leaf_size[3];  // full leaf box size per axis
  double k0;
  int nt, np;
  float2* out;
  int precision;
  // quad-symmetric BF16x3 kernel where eligible (np even): 1 (groups of 64
  // direction quads per block chosen by the launcher), 2 (one group per
  // block), 3 (all of a leaf's groups in one block); 0: plain kernels
  int quad;
  // 1: F = 6 S fields (S solves' E x, y, z, H x, y, z) are summed and the
  // pattern receives the two transverse components of each solve instead,
  // out [leaf][2 S][nt][np] as launch_project_transverse forms them
  // (leaf_fuses_transverse tells which launches can)
  int transverse = 0;
  float z0 = 0.f;
};
hipError_t launch_leaf(const LeafArgs& a, hipStream_t st);
bool leaf_fuses_transverse(const LeafArgs& a, bool bf16);
// MFMA variants (F <= 48 fields per launch): split BF16 x3 or FP32 MFMA.
hipError_t launch_leaf_mfma(const LeafArgs& a, bool bf16, hipStream_t st);

}  // namespace wf
