// wf_leaf_mfma.hip -- leaf radiation sum on the MFMA units.
//
// Per leaf L the pattern is the complex GEMM Y[d][f] = sum_n P[d][n] C[n][f],
// P[d][n] = exp(+i k s_d.(x_n - c_L)) (util/fields.h:8-38 summed per leaf as
// in treePO.cu:652-670).  Realified and transposed so that output lanes run
// along directions (coalesced stores):
//
//   Y^T[c][d] = sum_k Bt[c][k] Pt[k][d],   k = 2n + {re, im}, c = 2f + {re, im}
//   Pt[2n][d] = cos, Pt[2n+1][d] = sin of k s_d.(x_n - c_L)
//   Bt[2f][2n] = Re C, Bt[2f][2n+1] = -Im C, Bt[2f+1][2n] = Im C, Bt[2f+1][2n+1] = Re C
//
// MFMA A operand = Bt (currents of 16 points x 8 fields, staged once per leaf
// chunk in LDS in the per-lane register layout), MFMA B operand = Pt (phases
// generated in registers: each lane owns one direction and four points),
// accumulated in FP32.
//   * BF16X3: v_mfma_f32_16x16x32_bf16 with hi/lo splits of both operands,
//     Y += Ahi Bhi + Ahi Blo + Alo Bhi (Henry, Tang & Heinecke, ARITH 2019);
//     relative error ~1e-5 per product sum.
//   * FP32:   v_mfma_f32_16x16x4_f32.
// Operand lane layouts (gfx950, wave64):
//   16x16x32 bf16: A lane l -> row l%16, k = 8(l/16)+j; B lane l -> col l%16,
//                  k = 8(l/16)+j; D lane l -> col l%16, rows 4(l/16)+j.
//   16x16x4 f32:   A lane l -> row l%16, k = l/16; B lane l -> col l%16,
//                  k = l/16; D as above.
// Phases: local coordinates scaled to revolutions, reduced by rint, then the
// hardware sin/cos (inputs in revolutions).
#include <algorithm>
#include <cmath>

#include "wf_common.h"
#include "wf_interp.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int kThreads = 256;  // 4 waves

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

__device__ inline void leaf_dir(int d, int nt, int np, float s[3]) {
  const int j = d / np, kk = d - j * np;
  const double th = -WF_PI / 2 + WF_PI * (j + 0.5) / nt;
  const double ph = WF_TWO_PI * kk / np;
  double st, ct, sp, cp;
  sincos(th, &st, &ct);
  sincos(ph, &sp, &cp);
  s[0] = (float)(ct * cp);
  s[1] = (float)(ct * sp);
  s[2] = (float)st;
}

__device__ inline float rev_reduce(float r) { return r - rintf(r); }

// lanes of one wave exchange values through LDS: the wave's LDS requests
// complete in issue order, so a compiler barrier orders the phases
__device__ __forceinline__ void leaf_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Stage the scaled local coordinates of one chunk of points (zero-padded).
__device__ inline void stage_points(const wf::LeafArgs& a, int base, int cnt, int npts,
                                    const double c[3], double k_rev, float4* spt) {
  for (int t = threadIdx.x; t < npts; t += kThreads) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (t < cnt) {
      const int i = base + t;
      v.x = (float)(((double)a.px[i] - c[0]) * k_rev);
      v.y = (float)(((double)a.py[i] - c[1]) * k_rev);
      v.z = (float)(((double)a.pz[i] - c[2]) * k_rev);
    }
    spt[t] = v;
  }
}

// ---------------------------------------------------------------- BF16 x3 --
template <int MT, int R>
__global__ __launch_bounds__(kThreads) void leaf_bf16x3_kernel(wf::LeafArgs a, int KS) {
  extern __shared__ float4 smem4[];
  float4* spt = smem4;                                          // KS*16 points
  bf16x8* scur = reinterpret_cast<bf16x8*>(smem4 + KS * 16);    // [ks][mt][hi,lo][64]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, col = lane & 15;
  const int D = a.nt * a.np;
  const int dtiles = (D + 15) / 16;
  const int per_block = 4 * R;
  const int nsuper = (dtiles + per_block - 1) / per_block;
  const int leaf = (int)(blockIdx.x / nsuper), sup = (int)(blockIdx.x % nsuper);
  const int b = a.leaf_begin[leaf], e = a.leaf_end[leaf];
  double c[3];
  leaf_centre(a.leaf_key[leaf], a.depth, a.box_lo, a.leaf_size, c);
  const double k_rev = a.k0 / WF_TWO_PI;
  float s[R][3];
  int dd[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    dd[r] = (sup * per_block + wave * R + r) * 16 + col;
    s[r][0] = s[r][1] = s[r][2] = 0.f;
    if (dd[r] < D) leaf_dir(dd[r], a.nt, a.np, s[r]);
  }
  floatx4 acc[R][MT];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[r][m] = floatx4{0.f, 0.f, 0.f, 0.f};
  const int chunk = KS * 16;
  for (int base = b; base < e; base += chunk) {
    const int cnt = min(chunk, e - base);
    const int ksteps = (cnt + 15) / 16;
    __syncthreads();
    stage_points(a, base, cnt, ksteps * 16, c, k_rev, spt);
    // currents in the MFMA A-operand lane layout, hi and lo parts
    for (int idx = threadIdx.x; idx < ksteps * MT * 64; idx += kThreads) {
      const int ln = idx & 63, mt = (idx >> 6) % MT, ks = idx / (64 * MT);
      const int cc = ln & 15, gg = ln >> 4;
      const int f = mt * 8 + (cc >> 1), part = cc & 1;
      bf16x8 hi, lo;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int p = ks * 16 + 4 * gg + q;
        float2 C = make_float2(0.f, 0.f);
        if (p < cnt && f < a.F) C = a.C[(int64_t)(base + p) * a.Ftot + a.f0 + f];
        const float vr = part ? C.y : C.x;   // row 2n   (multiplies cos)
        const float vi = part ? C.x : -C.y;  // row 2n+1 (multiplies sin)
        const __bf16 hr = (__bf16)vr, hi_ = (__bf16)vi;
        hi[2 * q] = hr;
        hi[2 * q + 1] = hi_;
        lo[2 * q] = (__bf16)(vr - (float)hr);
        lo[2 * q + 1] = (__bf16)(vi - (float)hi_);
      }
      scur[(ks * MT + mt) * 128 + ln] = hi;
      scur[(ks * MT + mt) * 128 + 64 + ln] = lo;
    }
    __syncthreads();
    for (int ks = 0; ks < ksteps; ++ks) {
      float4 pt[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) pt[q] = spt[ks * 16 + 4 * g + q];
      bf16x8 ahi[MT], alo[MT];
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        ahi[m] = scur[(ks * MT + m) * 128 + lane];
        alo[m] = scur[(ks * MT + m) * 128 + 64 + lane];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        bf16x8 phi, plo;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float f = rev_reduce(s[r][0] * pt[q].x + s[r][1] * pt[q].y + s[r][2] * pt[q].z);
          const float cs = __builtin_amdgcn_cosf(f), sn = __builtin_amdgcn_sinf(f);
          const __bf16 hc = (__bf16)cs, hs = (__bf16)sn;
          phi[2 * q] = hc;
          phi[2 * q + 1] = hs;
          plo[2 * q] = (__bf16)(cs - (float)hc);
          plo[2 * q + 1] = (__bf16)(sn - (float)hs);
        }
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          acc[r][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[m], phi, acc[r][m], 0, 0, 0);
          acc[r][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[m], plo, acc[r][m], 0, 0, 0);
          acc[r][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[m], phi, acc[r][m], 0, 0, 0);
        }
      }
    }
  }
  // D lane l: column = direction dd, rows 4g + j -> fields mt*8 + 2g + j/2, part j%2
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (dd[r] >= D) continue;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int f = m * 8 + 2 * g + h;
        if (f < a.F)
          a.out[((int64_t)leaf * a.F + f) * D + dd[r]] =
              make_float2(acc[r][m][2 * h], acc[r][m][2 * h + 1]);
      }
    }
  }
}

// ------------------------------------------------------- BF16 x3, quads --
// The half-sample theta grid is symmetric (theta_{nt-1-j} = -theta_j) and
// np is even (phi_{k+np/2} = phi_k + pi), so directions come in quadruples
//   r0 = (j, k), r1 = (j, k + np/2), r2 = (nt-1-j, k), r3 = (nt-1-j, k + np/2)
// with phases A + B, -A + B, A - B, -A - B, A = k cos(theta)(x cos phi + y sin phi),
// B = k z sin(theta).  Two exponentials per point, e^{i(A+B)} and e^{i(A-B)},
// give all four (the other two are their conjugates).  Lane l owns quad
// tile*16 + l%16; the four directions are its four MFMA B operands.
// With RND a block takes G consecutive groups of 64 quads of one leaf (16
// per wave), one group per round; when the leaf's points fit one staging
// chunk and the epilogue leaves the LDS alone, the chunk staged in the first
// round serves all G rounds.  Without RND, G = 1.
template <int MT, bool RND>
__global__ __launch_bounds__(kThreads) void leaf_bf16x3_quad_kernel(wf::LeafArgs a, int KS, int RB,
                                                                    int G_) {
  extern __shared__ float4 smem4[];
  float4* spt = smem4;
  bf16x8* scur = reinterpret_cast<bf16x8*>(smem4 + KS * 16);
  // e^{iB} of every (direction row of this block, staged point): B = k s_z z
  // depends on the direction only through its row, so it is formed once per
  // (row, point) here instead of once per (direction, point) in the lanes
  float2* seb = reinterpret_cast<float2*>(scur + KS * MT * 128);  // [RB][KS*16]
  float* sbz = reinterpret_cast<float*>(seb + RB * KS * 16);       // [RB] sin(theta_row)
  const int wave = threadIdx.x >> 6;
  const int D = a.nt * a.np;
  const int nth = (a.nt + 1) / 2, nph = a.np / 2;
  const bool perm = (MT % 3 == 0) && a.transverse;  // solve-per-lane-group field order
  const int nquad = nth * nph;
  const int qtiles = (nquad + 15) / 16;
  const int nsuper = (qtiles + 3) / 4;
  const int G = RND ? G_ : 1;
  const int nsg = (nsuper + G - 1) / G;  // blocks per leaf
  const int leaf = (int)(blockIdx.x / nsg), sg = (int)(blockIdx.x % nsg);
  const int b = a.leaf_begin[leaf], e = a.leaf_end[leaf];
  double c[3];
  leaf_centre(a.leaf_key[leaf], a.depth, a.box_lo, a.leaf_size, c);
  const double k_rev = a.k0 / WF_TWO_PI;
  // rows of the block's G * 64 quads: jlo .. jlo + nrow - 1 (nrow <= RB, launcher)
  const int q0 = sg * G * 64;
  const int jlo = q0 / nph;
  const int nrow = min(nth - 1, (q0 + G * 64 - 1) / nph) - jlo + 1;
  // direction cosines in FP32 (sincospi of theta/pi, phi/pi): an error of an
  // ulp in s moves a leaf phase by ~1e-7 revolutions
  const float inv_nt = 1.f / (float)a.nt, inv_np = 2.f / (float)a.np;
  for (int r = threadIdx.x; r < nrow; r += kThreads) {
    float st_, ct_;
    sincospif(((float)(jlo + r) + 0.5f) * inv_nt - 0.5f, &st_, &ct_);
    sbz[r] = st_;
  }
  const int chunk = KS * 16;
  // the chunk staged in the first round serves every round: one chunk, and
  // no epilogue through the LDS
  const bool keep = (e - b) <= chunk && (!a.transverse || MT % 3 == 0);
  for (int rd = 0; rd < G; ++rd) {
    const int sup = sg * G + rd;
    if (sup >= nsuper) break;  // uniform per block
    // the lane index through an empty asm: the round's lane-derived LDS
    // addresses are formed in the round instead of being held across rounds
    int lr = (int)threadIdx.x & 63;
    if constexpr (RND) __asm__ volatile("" : "+v"(lr));
    const int lane = lr, g = lr >> 4, col = lr & 15;
    const int quad = (sup * 4 + wave) * 16 + col;
    const bool qvalid = quad < nquad;
    const bool wave_live = (sup * 4 + wave) * 16 < nquad;  // uniform per wave
    const int j = qvalid ? quad / nph : 0, k = qvalid ? quad - (quad / nph) * nph : 0;
    const int jrow = qvalid ? j - jlo : 0;
    const bool mirror_row = (a.nt - 1 - j) != j;
    int dd[4];
    dd[0] = j * a.np + k;
    dd[1] = j * a.np + k + nph;
    dd[2] = (a.nt - 1 - j) * a.np + k;
    dd[3] = (a.nt - 1 - j) * a.np + k + nph;
    // A = ax x + ay y (revolutions per metre scaled); B = s_z z comes from seb
    float ax = 0.f, ay = 0.f;
    if (qvalid) {
      float st, ct, sp, cp;
      sincospif(((float)j + 0.5f) * inv_nt - 0.5f, &st, &ct);
      sincospif((float)k * inv_np, &sp, &cp);
      ax = ct * cp;
      ay = ct * sp;
    }
    floatx4 acc[4][MT];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int m = 0; m < MT; ++m) acc[r][m] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int base = b; base < e; base += chunk) {
      const int cnt = min(chunk, e - base);
      const int ksteps = (cnt + 15) / 16;
      if (rd == 0 || !keep) {  // uniform per block
        __syncthreads();
        stage_points(a, base, cnt, ksteps * 16, c, k_rev, spt);
        // the same float z the staging forms, then the lanes' B phase and sincos
        for (int t = threadIdx.x; t < nrow * ksteps * 16; t += kThreads) {
          const int r = t / (ksteps * 16), p = t - r * (ksteps * 16);
          const float z = (p < cnt) ? (float)(((double)a.pz[base + p] - c[2]) * k_rev) : 0.f;
          const float fb = rev_reduce(sbz[r] * z);
          seb[r * chunk + p] = make_float2(__builtin_amdgcn_cosf(fb), __builtin_amdgcn_sinf(fb));
        }
        for (int idx = threadIdx.x; idx < ksteps * MT * 64; idx += kThreads) {
          const int ln = idx & 63, mt = (idx >> 6) % MT, ks = idx / (64 * MT);
          const int cc = ln & 15, gg = ln >> 4;
          // transverse passes of 3 or 6 tiles order the fields so that lane group
          // g of tiles 3q .. 3q + 2 holds the six fields of solve g + 4q
          const int fl = cc >> 1;
          const int f = perm ? 6 * ((fl >> 1) + 4 * (mt / 3)) + 2 * (mt % 3) + (fl & 1)
                             : mt * 8 + fl;
          const int part = cc & 1;
          bf16x8 hi, lo;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int p = ks * 16 + 4 * gg + q;
            float2 C = make_float2(0.f, 0.f);
            if (p < cnt && f < a.F) C = a.C[(int64_t)(base + p) * a.Ftot + a.f0 + f];
            const float vr = part ? C.y : C.x;
            const float vi = part ? C.x : -C.y;
            const __bf16 hr = (__bf16)vr, hi_ = (__bf16)vi;
            hi[2 * q] = hr;
            hi[2 * q + 1] = hi_;
            lo[2 * q] = (__bf16)(vr - (float)hr);
            lo[2 * q + 1] = (__bf16)(vi - (float)hi_);
          }
          scur[(ks * MT + mt) * 128 + ln] = hi;
          scur[(ks * MT + mt) * 128 + 64 + ln] = lo;
        }
        __syncthreads();
      }
      // a wave whose 16 quads all lie past the grid only helps stage
      for (int ks = 0; ks < (wave_live ? ksteps : 0); ++ks) {
        bf16x8 h0, l0, h2, l2;  // phases of r0 (A+B) and r2 (A-B)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float4 pt = spt[ks * 16 + 4 * g + q];
          const float fa = rev_reduce(ax * pt.x + ay * pt.y);
          const float cA = __builtin_amdgcn_cosf(fa), sA = __builtin_amdgcn_sinf(fa);
          const float2 eB = seb[jrow * chunk + ks * 16 + 4 * g + q];
          const float cB = eB.x, sB = eB.y;
          const float cc_ = cA * cB, ss = sA * sB, sc = sA * cB, cs = cA * sB;
          const float p0c = cc_ - ss, p0s = sc + cs;  // e^{i(A+B)}
          const float p2c = cc_ + ss, p2s = sc - cs;  // e^{i(A-B)}
          const __bf16 h0c = (__bf16)p0c, h0s = (__bf16)p0s, h2c = (__bf16)p2c, h2s = (__bf16)p2s;
          h0[2 * q] = h0c;
          h0[2 * q + 1] = h0s;
          h2[2 * q] = h2c;
          h2[2 * q + 1] = h2s;
          l0[2 * q] = (__bf16)(p0c - (float)h0c);
          l0[2 * q + 1] = (__bf16)(p0s - (float)h0s);
          l2[2 * q] = (__bf16)(p2c - (float)h2c);
          l2[2 * q + 1] = (__bf16)(p2s - (float)h2s);
        }
        // r1 = conj(r2), r3 = conj(r0): negate the sine entries (exact in bf16)
        bf16x8 h1 = h2, l1 = l2, h3 = h0, l3 = l0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          h1[2 * q + 1] = -h2[2 * q + 1];
          l1[2 * q + 1] = -l2[2 * q + 1];
          h3[2 * q + 1] = -h0[2 * q + 1];
          l3[2 * q + 1] = -l0[2 * q + 1];
        }
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          const bf16x8 ahi = scur[(ks * MT + m) * 128 + lane];
          const bf16x8 alo = scur[(ks * MT + m) * 128 + 64 + lane];
          acc[0][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, h0, acc[0][m], 0, 0, 0);
          acc[0][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, l0, acc[0][m], 0, 0, 0);
          acc[0][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, h0, acc[0][m], 0, 0, 0);
          acc[1][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, h1, acc[1][m], 0, 0, 0);
          acc[1][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, l1, acc[1][m], 0, 0, 0);
          acc[1][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, h1, acc[1][m], 0, 0, 0);
          acc[2][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, h2, acc[2][m], 0, 0, 0);
          acc[2][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, l2, acc[2][m], 0, 0, 0);
          acc[2][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, h2, acc[2][m], 0, 0, 0);
          acc[3][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, h3, acc[3][m], 0, 0, 0);
          acc[3][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi, l3, acc[3][m], 0, 0, 0);
          acc[3][m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo, h3, acc[3][m], 0, 0, 0);
        }
      }
    }
    if (a.transverse) {
      // Two transverse components per solve (F = 6 S fields): lane (col, g)
      // stages its fields f = 8 m + 2 g + h of direction r in the wave's LDS
      // slice [f][col], then lane (col, g) projects solves g and g + 4 onto the unit
      // vectors t (theta, latitude) and p (phi) of the direction:
      //   a = t.E - z0 p.H,  b = p.E + z0 t.H.
      // The quadruple's directions flip sin(theta) (r2, r3) and cos/sin(phi)
      // (r1, r3).
      const int S = a.F / 6;
      float sth = 0.f, cth = 0.f, sph = 0.f, cph = 0.f;
      if (qvalid) {
        sincospif(((float)j + 0.5f) * inv_nt - 0.5f, &sth, &cth);
        sincospif((float)k * inv_np, &sph, &cph);
      }
      const float z0 = a.z0;
      if constexpr (MT % 3 == 0) {
        // permuted fields: tile 3q + i, half h holds component 2 i + h of solve
        // g + 4 q (E x, y, z, H x, y, z), projected in registers
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (!qvalid) break;
          if (r >= 2 && !mirror_row) break;
          const float sr = (r & 2) ? -sth : sth;
          const float pc = (r & 1) ? -cph : cph, ps = (r & 1) ? -sph : sph;
          const float t[3] = {-sr * pc, -sr * ps, cth}, p[3] = {-ps, pc, 0.f};
#pragma unroll
          for (int q = 0; q < MT / 3; ++q) {
            const int sv = g + 4 * q;
            if (sv >= S) continue;
            float2 fld[6];
#pragma unroll
            for (int c = 0; c < 6; ++c)
              fld[c] = make_float2(acc[r][3 * q + c / 2][2 * (c & 1)],
                                   acc[r][3 * q + c / 2][2 * (c & 1) + 1]);
            float2 av = make_float2(0.f, 0.f), bv = make_float2(0.f, 0.f);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              const float2 e = fld[c], h = fld[3 + c];
              av.x += t[c] * e.x - z0 * p[c] * h.x;
              av.y += t[c] * e.y - z0 * p[c] * h.y;
              bv.x += p[c] * e.x + z0 * t[c] * h.x;
              bv.y += p[c] * e.y + z0 * t[c] * h.y;
            }
            float2* o = a.out + ((int64_t)leaf * 2 * S + 2 * sv) * D + dd[r];
            o[0] = av;
            o[D] = bv;
          }
        }
      } else {
        __syncthreads();  // every wave done with the staged points and currents
        float2* ep = reinterpret_cast<float2*>(smem4) + wave * (MT * 8 * 16);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
#pragma unroll
          for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int hh = 0; hh < 2; ++hh)
              ep[(m * 8 + 2 * g + hh) * 16 + col] =
                  make_float2(acc[r][m][2 * hh], acc[r][m][2 * hh + 1]);
          leaf_wave_sync();
          if (qvalid && (r < 2 || mirror_row)) {
            const float sr = (r & 2) ? -sth : sth;
            const float pc = (r & 1) ? -cph : cph, ps = (r & 1) ? -sph : sph;
            const float t[3] = {-sr * pc, -sr * ps, cth}, p[3] = {-ps, pc, 0.f};
            for (int sv = g; sv < S; sv += 4) {  // solves g, g + 4 (S <= 8)
              const float2* fs = ep + (6 * sv) * 16 + col;
              float2 av = make_float2(0.f, 0.f), bv = make_float2(0.f, 0.f);
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const float2 e = fs[c * 16], h = fs[(3 + c) * 16];
                av.x += t[c] * e.x - z0 * p[c] * h.x;
                av.y += t[c] * e.y - z0 * p[c] * h.y;
                bv.x += p[c] * e.x + z0 * t[c] * h.x;
                bv.y += p[c] * e.y + z0 * t[c] * h.y;
              }
              float2* o = a.out + ((int64_t)leaf * 2 * S + 2 * sv) * D + dd[r];
              o[0] = av;
              o[D] = bv;
            }
          }
          leaf_wave_sync();
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // middle row of an odd grid: r2, r3 repeat r0, r1
        if (!qvalid || (r >= 2 && !mirror_row)) break;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
#pragma unroll
          for (int hh = 0; hh < 2; ++hh) {
            const int f = m * 8 + 2 * g + hh;
            if (f < a.F)
              a.out[((int64_t)leaf * a.F + f) * D + dd[r]] =
                  make_float2(acc[r][m][2 * hh], acc[r][m][2 * hh + 1]);
          }
        }
      }
    }
  }  // rounds
}

// ------------------------------------------------------------------- FP32 --
template <int MT, int R>
__global__ __launch_bounds__(kThreads) void leaf_f32_kernel(wf::LeafArgs a, int KS) {
  extern __shared__ float4 smem4[];
  float4* spt = smem4;                                      // KS*16 points
  float* scur = reinterpret_cast<float*>(smem4 + KS * 16);  // [kstep(2 pts)][mt][64]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, col = lane & 15;
  const int D = a.nt * a.np;
  const int dtiles = (D + 15) / 16;
  const int per_block = 4 * R;
  const int nsuper = (dtiles + per_block - 1) / per_block;
  const int leaf = (int)(blockIdx.x / nsuper), sup = (int)(blockIdx.x % nsuper);
  const int b = a.leaf_begin[leaf], e = a.leaf_end[leaf];
  double c[3];
  leaf_centre(a.leaf_key[leaf], a.depth, a.box_lo, a.leaf_size, c);
  const double k_rev = a.k0 / WF_TWO_PI;
  float s[R][3];
  int dd[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    dd[r] = (sup * per_block + wave * R + r) * 16 + col;
    s[r][0] = s[r][1] = s[r][2] = 0.f;
    if (dd[r] < D) leaf_dir(dd[r], a.nt, a.np, s[r]);
  }
  floatx4 acc[R][MT];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int m = 0; m < MT; ++m) acc[r][m] = floatx4{0.f, 0.f, 0.f, 0.f};
  // lane's k = g: point 2*step + g/2, component g%2 (0: cos row, 1: sin row)
  const float quarter = (g & 1) ? 0.f : 0.25f;  // sin(2 pi (f + 1/4)) = cos(2 pi f)
  const int chunk = KS * 16;
  for (int base = b; base < e; base += chunk) {
    const int cnt = min(chunk, e - base);
    const int ksteps = (cnt + 15) / 16;
    const int nsub = ksteps * 8;  // MFMA k-steps of two points
    __syncthreads();
    stage_points(a, base, cnt, ksteps * 16, c, k_rev, spt);
    for (int idx = threadIdx.x; idx < nsub * MT * 64; idx += kThreads) {
      const int ln = idx & 63, mt = (idx >> 6) % MT, st = idx / (64 * MT);
      const int cc = ln & 15, gg = ln >> 4;
      const int f = mt * 8 + (cc >> 1), part = cc & 1;
      const int p = 2 * st + (gg >> 1);
      float2 C = make_float2(0.f, 0.f);
      if (p < cnt && f < a.F) C = a.C[(int64_t)(base + p) * a.Ftot + a.f0 + f];
      // Bt[c][k]: k even (cos row) -> part ? Im : Re ; k odd (sin row) -> part ? Re : -Im
      scur[(st * MT + mt) * 64 + ln] = (gg & 1) ? (part ? C.x : -C.y) : (part ? C.y : C.x);
    }
    __syncthreads();
    for (int st = 0; st < nsub; ++st) {
      const float4 pt = spt[2 * st + (g >> 1)];
      float am[MT];
#pragma unroll
      for (int m = 0; m < MT; ++m) am[m] = scur[(st * MT + m) * 64 + lane];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const float f = rev_reduce(s[r][0] * pt.x + s[r][1] * pt.y + s[r][2] * pt.z) + quarter;
        const float ph = __builtin_amdgcn_sinf(f);
#pragma unroll
        for (int m = 0; m < MT; ++m)
          acc[r][m] = __builtin_amdgcn_mfma_f32_16x16x4f32(am[m], ph, acc[r][m], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (dd[r] >= D) continue;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int f = m * 8 + 2 * g + h;
        if (f < a.F)
          a.out[((int64_t)leaf * a.F + f) * D + dd[r]] =
              make_float2(acc[r][m][2 * h], acc[r][m][2 * h + 1]);
      }
    }
  }
}

template <int MT>
hipError_t launch_mt(const wf::LeafArgs& a, bool bf16, hipStream_t st) {
  constexpr int R = (MT <= 2) ? 4 : (MT <= 4 ? 2 : 1);
  const int D = a.nt * a.np;
  if (bf16 && a.quad && (a.np % 2) == 0) {
    const int nquad = ((a.nt + 1) / 2) * (a.np / 2);
    const int nsuper = ((nquad + 15) / 16 + 3) / 4;
    const size_t per_ks = (size_t)MT * 128 * sizeof(bf16x8);
    int KS = (int)((48 * 1024) / per_ks);
    KS = KS < 1 ? 1 : (KS > 8 ? 8 : KS);
    // staged points per chunk: about 1.25x the mean leaf, so small leaves do
    // not reserve LDS (and with it blocks per CU) for points they never have
    const double mean = a.nleaves > 0 ? (double)a.n / a.nleaves : 0.0;
    const int ks_mean = (int)std::ceil(1.25 * mean / 16.0);
    if (ks_mean >= 2 && ks_mean < KS) KS = ks_mean;
    if (a.transverse && a.F % 6 != 0) return hipErrorInvalidValue;
    // groups of 64 quads per block (rounds): when the typical leaf fits one
    // staging chunk and the epilogue keeps the LDS, a block runs all groups
    // of its leaf on one staging, as long as the grid keeps >= 2048 blocks
    // (not for 2 or 4 tiles, whose round form needs registers for one wave
    // per SIMD less)
    // (quad 2: one group per block; quad 3: all of a leaf's groups in one
    // block, whatever the grid)
    const bool reuse = (!a.transverse || MT % 3 == 0) && MT != 2 && MT != 4 &&
                       1.25 * mean <= KS * 16.0;
    int G = 1;
    if (a.quad == 3) {
      G = nsuper;
    } else if (a.quad == 1 && reuse) {
      G = nsuper;
      while (G > 1 && (int64_t)a.nleaves * ((nsuper + G - 1) / G) < 2048) G = (G + 1) / 2;
    }
    const int nph = a.np / 2, nth = (a.nt + 1) / 2;
    // LDS of a block: up to 80 KB, two blocks per CU
    constexpr size_t kMaxLds = 80 * 1024;
    size_t sh = 0;
    int RB = 0, KSb = KS;
    for (;;) {
      // rounds: a leaf that fits one chunk is staged once for all its groups,
      // so the chunk holds 1.6x the mean leaf, nearly every leaf (663k-point
      // fiber: 98 % of its leaves have at most 80 points, 70 % at most 64)
      KSb = KS;
      if (G > 1) {
        const int ks_r = std::min({(int)std::ceil(1.6 * mean / 16.0),
                                   (int)((64 * 1024) / per_ks), 8});
        KSb = std::max(KS, ks_r);
      }
      // direction rows one block's G * 64 consecutive quads can touch
      RB = std::min(nth, (G * 64 - 1) / nph + 2);
      sh = (size_t)KSb * 16 * sizeof(float4) + (size_t)KSb * per_ks +
           (size_t)RB * KSb * 16 * sizeof(float2) + (size_t)RB * sizeof(float);
      if (a.transverse) {  // the epilogue's per-wave [field][16 directions] slices
        const size_t ep = (size_t)4 * MT * 8 * 16 * sizeof(float2);
        sh = sh > ep ? sh : ep;
      }
      if (sh <= kMaxLds || G == 1) break;
      G = 1;
    }
    if (sh > kMaxLds) return hipErrorInvalidValue;
    const int64_t nblk = (int64_t)((nsuper + G - 1) / G) * a.nleaves;
    if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
    const void* kern = G > 1 ? (const void*)leaf_bf16x3_quad_kernel<MT, true>
                             : (const void*)leaf_bf16x3_quad_kernel<MT, false>;
    if (sh > 64 * 1024) {  // above the default dynamic LDS limit
      hipError_t e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
      if (e != hipSuccess) return e;
    }
    if (G > 1)
      hipLaunchKernelGGL((leaf_bf16x3_quad_kernel<MT, true>), dim3((unsigned)nblk), dim3(kThreads),
                         sh, st, a, KSb, RB, G);
    else
      hipLaunchKernelGGL((leaf_bf16x3_quad_kernel<MT, false>), dim3((unsigned)nblk), dim3(kThreads),
                         sh, st, a, KSb, RB, 1);
    return hipGetLastError();
  }
  if (a.transverse) return hipErrorInvalidValue;  // the quad kernel only
  const int dtiles = (D + 15) / 16;
  const int nsuper = (dtiles + 4 * R - 1) / (4 * R);
  const int64_t nblk = (int64_t)nsuper * a.nleaves;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  const size_t per_ks = bf16 ? (size_t)MT * 128 * sizeof(bf16x8) : (size_t)8 * MT * 64 * sizeof(float);
  int KS = (int)((48 * 1024) / per_ks);
  KS = KS < 1 ? 1 : (KS > 8 ? 8 : KS);
  const size_t sh = (size_t)KS * 16 * sizeof(float4) + (size_t)KS * per_ks;
  if (bf16)
    hipLaunchKernelGGL((leaf_bf16x3_kernel<MT, R>), dim3((unsigned)nblk), dim3(kThreads), sh, st, a, KS);
  else
    hipLaunchKernelGGL((leaf_f32_kernel<MT, R>), dim3((unsigned)nblk), dim3(kThreads), sh, st, a, KS);
  return hipGetLastError();
}

}  // namespace

namespace wf {

bool leaf_fuses_transverse(const LeafArgs& a, bool bf16) {
  return bf16 && a.quad && a.F % 6 == 0 && (a.F + 7) / 8 <= 6 && (a.np % 2) == 0 &&
         a.f0 == 0 && a.F == a.Ftot;
}

// F complex fields handled per launch: up to 48 (MT = 6 tiles of 8 fields).
hipError_t launch_leaf_mfma(const LeafArgs& a, bool bf16, hipStream_t st) {
  if (a.nleaves <= 0) return hipSuccess;
  const int MT = (a.F + 7) / 8;
  switch (MT) {
    case 1: return launch_mt<1>(a, bf16, st);
    case 2: return launch_mt<2>(a, bf16, st);
    case 3: return launch_mt<3>(a, bf16, st);
    case 4: return launch_mt<4>(a, bf16, st);
    case 5: return launch_mt<5>(a, bf16, st);
    case 6: return launch_mt<6>(a, bf16, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace wf
