// wf_agg_mfma.hip -- one level transition of the tree aggregation as dense
// interpolation GEMMs on the MFMA units.
//
// The reference interpolates with cuFFT plans per child and component
// (interpolation.h:11-92), translates (fields.h:92-103) and sums children
// (fields.h:153-170).  Resampling a band-limited periodic sequence of length
// nc onto np points is a fixed real linear map: the child's trigonometric
// interpolant evaluated at the parent's sample positions,
//
//   M[q][j] = D_nc(x_q - 2 pi j / nc),
//   D_n(d)  = (1/n) sum_{f in band(n)} w_f e^{i f d}
//           = sin(n d/2) / (n sin(d/2))          (n odd)
//           = sin(n d/2) cot(d/2) / n            (n even; Nyquist halves)
//
// which is exactly what FFT -> pad/fold -> inverse FFT computes.  So a level
// transition is two GEMMs with a small shared real matrix:
//
//   phi:   G[cf][theta][q] = sum_j X[cf][theta][j] Mphi[q][j]          (q < npp)
//   theta: U[e][k]         = sum_j' Mth[e][j'] S[j'][k]                (e < 2 ntp)
//          S[j'][k] = G[j'][k] (j' < ntc), G[2 ntc - 1 - j'][k + npp/2] (j' >= ntc)
//          (the column pair phi_k, phi_k + pi is one great circle), then
//          parent[e][k] (e < ntp) / parent[mirror - e][k + npp/2] += T U.
//
// Both run on v_mfma_f32_16x16x32_bf16 with split-BF16 x3 operands (hi*hi +
// hi*lo + lo*hi), FP32 accumulation; the root level accumulates the child sum
// in FP64.  The matrices are split on the host; the data are split once, by
// the phi pass, which stores for every column pair k one row holding column
// k and then column k + npp/2, both theta-ascending and padded to ceil8(ntc),
// as groups of 8 samples in 4 BF16 planes (re hi, re lo, im hi, im lo).  The
// theta matrix's columns are permuted to that order (build_theta_matrix_gc),
// so the theta pass copies 64 B groups straight into LDS.  Each pass keeps
// its matrix fragments in registers for a whole persistent block.
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "wf_common.h"
#include "wf_interp.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

__device__ inline void split8(const float v[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const __bf16 h = (__bf16)v[j];
    hi[j] = h;
    lo[j] = (__bf16)(v[j] - (float)h);
  }
}

__device__ inline floatx4 mfma3(const bf16x8& ah, const bf16x8& al, const bf16x8& bh,
                                const bf16x8& bl, floatx4 c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  return c;
}

typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;

// phi pass, matrix-stationary.  Rows are virtual: plane cf = (node, field)
// owns ntc8 = ceil8(ntc) consecutive rows (theta), the ones with theta >= ntc
// read as zero, so every aligned group of 8 rows lies in one plane and starts
// at a theta that is a multiple of 8.  Block = 4 waves side by side in q
// (wave: QT 16-column tiles, block: 64 QT columns); every wave keeps its
// matrix fragments (QT tiles x NKS K steps, hi and lo) in registers.  The
// block walks RT-row tiles (32, or 16 to halve the LDS so that two blocks
// share a CU) with stride ng; each tile (RT rows x npc,
// contiguous in X within a plane) is loaded coalesced, split into BF16 hi/lo
// and staged in LDS [buffer][k step][plane][k group][row] (MFMA lanes read
// consecutive 16 B), double buffered with one barrier per tile and the next
// D tiles in flight in registers.  Output: the D lane holds rows 4g..4g+3 of
// column q, i.e. half of an 8-sample group; it splits them into BF16 hi/lo
// and writes 8 B into each of the group's 4 planes (re hi, re lo, im hi,
// im lo) of the pre-split great-circle row of q.
template <int NKS, int QT, int D, int RT>
__global__ __launch_bounds__(256) void phi_tile_kernel(const float2* __restrict__ X,
                                                       bf16x4* __restrict__ GT, int planes,
                                                       int ntc, int npc, int npp,
                                                       const bf16x8* __restrict__ Mhi,
                                                       const bf16x8* __restrict__ Mlo, int kpad,
                                                       int mrows, int nqb, int ng, int64_t used) {
  constexpr int MT = RT / 16;          // 16-row MFMA tiles per row tile
  constexpr int NITEM = RT * NKS * 4;  // (row, 8-sample group) items per tile
  constexpr int LI = (NITEM + 255) / 256;
  __shared__ bf16x8 As[2][NKS][4][4][RT];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c16 = lane & 15, g = lane >> 4;
  const int64_t nb = gridDim.x, per = (nb + 7) / 8;
  const int64_t w = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (w >= used) return;
  const int qb = (int)(w % nqb);
  const int grp = (int)(w / nqb);
  const int ntc8 = (ntc + 7) & ~7;
  const int rows = planes * ntc8;  // virtual rows
  const int nrt = (rows + RT - 1) / RT;
  if (grp >= nrt) return;
  const int nks = (npc + 31) >> 5;
  const int half = npp / 2;
  const int ngrp = ntc8 >> 2;  // 8-sample groups per great-circle row (2 ntc8 / 8)
  const int q0 = qb * 64 * QT + wave * 16 * QT;
  // resident matrix fragments: B lane (col q = c16, k group g)
  bf16x8 bh[QT][NKS], bl[QT][NKS];
#pragma unroll
  for (int n = 0; n < QT; ++n) {
    const int q = q0 + 16 * n + c16;
    const int qq = q < mrows ? q : 0;
    const bf16x8* ph = Mhi + ((qq * kpad) >> 3);
    const bf16x8* pl = Mlo + ((qq * kpad) >> 3);
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int kq = ks < nks ? ks * 4 + g : 0;
      bh[n][ks] = ph[kq];
      bl[n][ks] = pl[kq];
    }
  }
  // load items: i -> row i % RT, (ks, g) = i / RT
  float2 R[D][LI][8];
  auto load = [&](int d, int tile) {
#pragma unroll
    for (int q = 0; q < LI; ++q) {
      const int i = threadIdx.x + 256 * q;
      const int rv = tile * RT + (i % RT);
      const int cf = (int)((uint32_t)rv / (uint32_t)ntc8), th = rv - cf * ntc8;
      const bool rok = i < NITEM && rv < rows && th < ntc;
      const float4* src =
          reinterpret_cast<const float4*>(X + (int64_t)(rok ? cf * ntc + th : 0) * npc);
      const int jb = (i / RT) * 8;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int jj = jb + 2 * j;  // npc even: a pair never straddles the row end
        const bool ok = rok && jj < npc;
        float4 v = src[ok ? (jj >> 1) : 0];
        if (!ok) v = make_float4(0.f, 0.f, 0.f, 0.f);
        R[d][q][2 * j] = make_float2(v.x, v.y);
        R[d][q][2 * j + 1] = make_float2(v.z, v.w);
      }
    }
  };
  auto stage = [&](int d, int buf) {
#pragma unroll
    for (int q = 0; q < LI; ++q) {
      const int i = threadIdx.x + 256 * q;
      if (i < NITEM) {
        const int row = i % RT, kg = i / RT;  // kg = ks * 4 + g
        float vr[8], vi[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          vr[j] = R[d][q][j].x;
          vi[j] = R[d][q][j].y;
        }
        bf16x8 a0, a1, a2, a3;
        split8(vr, a0, a1);
        split8(vi, a2, a3);
        As[buf][kg >> 2][0][kg & 3][row] = a0;
        As[buf][kg >> 2][1][kg & 3][row] = a1;
        As[buf][kg >> 2][2][kg & 3][row] = a2;
        As[buf][kg >> 2][3][kg & 3][row] = a3;
      }
    }
  };
  int tile = grp, ltile = grp;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (ltile < nrt) load(d, ltile);
    ltile += ng;
  }
  int buf = 0;
  bool live = true;
  while (live) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (!live) break;
      stage(d, buf);
      __syncthreads();
      if (ltile < nrt) load(d, ltile);
      ltile += ng;
      floatx4 accr[MT][QT], acci[MT][QT];
#pragma unroll
      for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < QT; ++n) {
          accr[m][n] = floatx4{0.f, 0.f, 0.f, 0.f};
          acci[m][n] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        if (ks < nks) {
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            const int row = 16 * m + c16;
            const bf16x8 arh = As[buf][ks][0][g][row], arl = As[buf][ks][1][g][row],
                         aih = As[buf][ks][2][g][row], ail = As[buf][ks][3][g][row];
#pragma unroll
            for (int n = 0; n < QT; ++n) {
              accr[m][n] = mfma3(arh, arl, bh[n][ks], bl[n][ks], accr[m][n]);
              acci[m][n] = mfma3(aih, ail, bh[n][ks], bl[n][ks], acci[m][n]);
            }
          }
        }
      }
      // D lane: column q, rows 4g..4g+3 of each M tile = half of one group
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int r4 = tile * RT + 16 * m + 4 * g;  // first of the lane's 4 rows
        if (r4 < rows) {
          const int cf = (int)((uint32_t)r4 / (uint32_t)ntc8), th = r4 - cf * ntc8;
#pragma unroll
          for (int n = 0; n < QT; ++n) {
            const int q = q0 + 16 * n + c16;
            if (q < npp) {
              const bool second = q >= half;
              const int k = second ? q - half : q;
              const int jg = ((second ? ntc8 : 0) + th) >> 3;
              bf16x4 rh, rl, ih, il;
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                const float vr = accr[m][n][j], vi = acci[m][n][j];
                const __bf16 hr = (__bf16)vr, hi_ = (__bf16)vi;
                rh[j] = hr;
                rl[j] = (__bf16)(vr - (float)hr);
                ih[j] = hi_;
                il[j] = (__bf16)(vi - (float)hi_);
              }
              // group (cf, k, jg): 4 planes x 2 halves of 4 samples
              bf16x4* o = GT + ((((int64_t)cf * half + k) * ngrp + jg) << 3) + (g & 1);
              o[0] = rh;
              o[2] = rl;
              o[4] = ih;
              o[6] = il;
            }
          }
        }
      }
      tile += ng;
      live = tile < nrt;
      buf ^= 1;
    }
  }
}

template <bool ROOT>
struct AccT;
template <>
struct AccT<false> {
  typedef float2 type;
};
template <>
struct AccT<true> {
  typedef double2 type;
};

struct ThetaK {
  const uint4* GT;  // [child][F][npp/2][n2c8 / 8][4 planes], pre-split great-circle rows
  int F, ntc, npp, ntp, n2c8;
  int nitems, ng, ngF, ngR;  // items (parent, field); stride ng = ngF F + ngR
  int neb, ncb, kpad, mrows, mirror;
  int64_t used;  // blocks with work
  const int32_t* cb;
  const int32_t* ce;
  const uint32_t* child_key;
  const float2* Text;
  int text_per_child;  // Text indexed by child instead of octant
  const bf16x8* Mhi;
  const bf16x8* Mlo;
  float2* out;
  double2* out_root;
  float dual_sign;  // -1 for transverse components: the pole's second sample
};

// theta pass + translation + child sum.  Block = 4 waves stacked along the
// great-circle rows e (BE = 64 ET rows) x BC = 16 CT column pairs k.  Each
// wave keeps its matrix fragments (ET row tiles x all NKS K steps, hi and
// lo) in registers for the block's lifetime, so the matrix is read once per
// block.  The block walks the (parent, field) items grp, grp + ng, ...; for
// every child of an item it stages the child's column tile (all K steps),
// split into BF16 hi/lo, in LDS ([buffer][k step][plane][column][k group]:
// staging writes and MFMA reads both cover consecutive 16 B, free of bank
// conflicts), double buffered with one barrier per
// child, while the data of the next D children are in flight in registers.
// The great-circle rows are contiguous in GT and already split, so each load
// item is one 64 B group (4 planes x 8 samples) copied to LDS unchanged.  The parent accumulates in registers in the D layout (lane:
// column k, rows 4g + j); row e of the extended grid feeds parent (e, k)
// when e < ntp and (mirror - e, k + npp/2) otherwise when that row is < ntp;
// with a sample at the pole (mirror = 2 ntp) row 0 also feeds (0, k +
// npp/2), kept in a second accumulator.  Lanes whose (e, k) lies outside the
// grid accumulate harmless values that are never stored.
template <int ET, int CT, int NKS, int D, bool ROOT, int MINB>
__global__ __launch_bounds__(256, MINB) void theta_mfma_kernel(const ThetaK P) {
  typedef typename AccT<ROOT>::type acc_t;
  constexpr int BE = 4 * ET * 16, BC = CT * 16;
  constexpr int NITEM = NKS * 4 * BC;  // 8-sample load items per child
  constexpr int LI = (NITEM + 255) / 256;
  __shared__ bf16x8 Bs[2][NKS][4][BC][4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c16 = lane & 15, g = lane >> 4;
  // XCD-aware mapping: blocks b = x (mod 8) run on XCD x and take a
  // contiguous range of work items
  const int64_t nb = gridDim.x, per = (nb + 7) / 8;
  const int64_t w = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (w >= P.used) return;  // whole block
  const int cbk = (int)(w % P.ncb);
  const int64_t t1 = w / P.ncb;
  const int eb = (int)(t1 % P.neb);
  const int grp = (int)(t1 / P.neb);
  if (grp >= P.nitems) return;  // whole block
  const int F = P.F, ntp = P.ntp, npp = P.npp, n2c8 = P.n2c8;
  const int half = npp / 2, n2p = 2 * ntp;
  const int nks = (n2c8 + 31) >> 5;
  const int e0 = eb * BE + wave * ET * 16;
  const int k0 = cbk * BC;
  const int i2_0 = P.mirror % n2p;
  const bool dual = (i2_0 < ntp) && e0 == 0 && g == 0;  // this lane holds row 0
  // resident matrix fragments
  bf16x8 ah[ET][NKS], al[ET][NKS];
#pragma unroll
  for (int a = 0; a < ET; ++a) {
    const int row = e0 + 16 * a + c16;
    const int rr = row < P.mrows ? row : 0;
    const bf16x8* ph = P.Mhi + ((rr * P.kpad) >> 3);
    const bf16x8* pl = P.Mlo + ((rr * P.kpad) >> 3);
#pragma unroll
    for (int ks = 0; ks < NKS; ++ks) {
      const int kq = ks < nks ? ks * 4 + g : 0;
      ah[a][ks] = ph[kq];
      al[a][ks] = pl[kq];
    }
  }
  // load items: i -> k group i & 3, column (i >> 2) % BC, K step (i >> 2) / BC
  int l_off[LI];
  bool l_ok[LI];
#pragma unroll
  for (int q = 0; q < LI; ++q) {
    const int i = threadIdx.x + 256 * q;
    const int lg = i & 3, lcol = (i >> 2) % BC, lks = (i >> 2) / BC;
    const int lk = k0 + lcol;
    const int jb = (lks * 4 + lg) * 8;
    l_ok[q] = i < NITEM && lk < half && jb < n2c8;
    l_off[q] = l_ok[q] ? (lk * n2c8 + jb) >> 1 : 0;  // 16 B units: 8 samples = 4 x 16 B
  }
  // translation-table offsets (clamped for lanes that are never stored)
  int tofs[ET][CT][4], tofs2[CT];
#pragma unroll
  for (int b = 0; b < CT; ++b) {
    const int col = k0 + 16 * b + c16;
    const bool cok = col < half;
#pragma unroll
    for (int a = 0; a < ET; ++a)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = e0 + 16 * a + 4 * g + j;
        tofs[a][b][j] = (cok && e < n2p) ? e * half + col : 0;
      }
    tofs2[b] = cok ? n2p * half + col : 0;
  }
  const int64_t slab = (int64_t)half * n2c8 / 2;  // 16 B units per (child, field)
  uint4 R[D][LI][4];
  int okey[D];
  auto load = [&](int d, int c, int f) {
    const uint4* S = P.GT + ((int64_t)c * F + f) * slab;
#pragma unroll
    for (int q = 0; q < LI; ++q) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        uint4 v = S[l_off[q] + j];
        if (!l_ok[q]) v = make_uint4(0u, 0u, 0u, 0u);
        R[d][q][j] = v;
      }
    }
    okey[d] = P.text_per_child ? c : (P.child_key ? (int)(P.child_key[c] & 7u) : 0);
  };
  auto stage = [&](int d, int buf) {
#pragma unroll
    for (int q = 0; q < LI; ++q) {
      const int i = threadIdx.x + 256 * q;
      if (i < NITEM) {
        const int lg = i & 3, lcol = (i >> 2) % BC, lks = (i >> 2) / BC;
#pragma unroll
        for (int pl = 0; pl < 4; ++pl)
          Bs[buf][lks][pl][lcol][lg] = *reinterpret_cast<const bf16x8*>(&R[d][q][pl]);
      }
    }
  };
  // (item, child) cursors: compute, and loads D children ahead
  auto advance = [&](int& xit, int& xp, int& xf, int& xc, int& xcend) {
    if (++xc >= xcend) {
      xit += P.ng;
      xf += P.ngR;
      xp += P.ngF;
      if (xf >= F) {
        xf -= F;
        ++xp;
      }
      if (xit < P.nitems) {
        xc = P.cb ? P.cb[xp] : xp;
        xcend = P.cb ? P.ce[xp] : xp + 1;
      }
    }
  };
  // does the compute cursor have a child after (xit, xc)?
  auto next_exists = [&](int xit, int xc, int xcend) {
    return xc + 1 < xcend || xit + P.ng < P.nitems;
  };
  int it = grp;
  int p = it / F, f = it - p * F;
  int c = P.cb ? P.cb[p] : p, cend = P.cb ? P.ce[p] : p + 1;
  int lit = it, lp = p, lf = f, lc = c, lcend = cend;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    if (lit < P.nitems) load(d, lc, lf);
    advance(lit, lp, lf, lc, lcend);
  }
  acc_t acc[ET][CT][4];
  acc_t acc2[CT];
#pragma unroll
  for (int a = 0; a < ET; ++a)
#pragma unroll
    for (int b = 0; b < CT; ++b)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[a][b][j].x = 0;
        acc[a][b][j].y = 0;
      }
#pragma unroll
  for (int b = 0; b < CT; ++b) {
    acc2[b].x = 0;
    acc2[b].y = 0;
  }
  // translation factors of the next child to be computed (unit if none)
  float2 tfn[ET][CT][4], tf2n[CT];
#pragma unroll
  for (int b = 0; b < CT; ++b) {
#pragma unroll
    for (int a = 0; a < ET; ++a)
#pragma unroll
      for (int j = 0; j < 4; ++j) tfn[a][b][j] = make_float2(1.f, 0.f);
    tf2n[b] = make_float2(1.f, 0.f);
  }
  auto load_t = [&](int oct) {
    const float2* To = P.Text + (int64_t)oct * (n2p + 1) * half;
#pragma unroll
    for (int b = 0; b < CT; ++b) {
#pragma unroll
      for (int a = 0; a < ET; ++a)
#pragma unroll
        for (int j = 0; j < 4; ++j) tfn[a][b][j] = To[tofs[a][b][j]];
      tf2n[b] = To[tofs2[b]];
    }
  };
  if (P.Text) load_t(okey[0]);
  // children still to stage after the current one (compute order)
  int buf = 0;
  bool live = true;
  while (live) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      if (!live) break;
      stage(d, buf);
      __syncthreads();
      const bool more_next = next_exists(it, c, cend);
      if (lit < P.nitems) load(d, lc, lf);
      advance(lit, lp, lf, lc, lcend);
      // translation factors: this child's were loaded one child ahead; issue
      // the next child's now (its octant is already in the queue)
      float2 tf[ET][CT][4], tf2[CT];
#pragma unroll
      for (int b = 0; b < CT; ++b) {
#pragma unroll
        for (int a = 0; a < ET; ++a)
#pragma unroll
          for (int j = 0; j < 4; ++j) tf[a][b][j] = tfn[a][b][j];
        tf2[b] = tf2n[b];
      }
      if (P.Text && more_next) load_t(okey[(d + 1) % D]);
      floatx4 ur[ET][CT], ui[ET][CT];
#pragma unroll
      for (int a = 0; a < ET; ++a)
#pragma unroll
        for (int b = 0; b < CT; ++b) {
          ur[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
          ui[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        if (ks < nks) {
#pragma unroll
          for (int b = 0; b < CT; ++b) {
            const int col = 16 * b + c16;
            const bf16x8 brh = Bs[buf][ks][0][col][g], brl = Bs[buf][ks][1][col][g],
                         bih = Bs[buf][ks][2][col][g], bil = Bs[buf][ks][3][col][g];
#pragma unroll
            for (int a = 0; a < ET; ++a) {
              ur[a][b] = mfma3(ah[a][ks], al[a][ks], brh, brl, ur[a][b]);
              ui[a][b] = mfma3(ah[a][ks], al[a][ks], bih, bil, ui[a][b]);
            }
          }
        }
      }
#pragma unroll
      for (int b = 0; b < CT; ++b)
#pragma unroll
        for (int a = 0; a < ET; ++a)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const float2 u = make_float2(ur[a][b][j], ui[a][b][j]);
            const float2 tu = cmul(tf[a][b][j], u);
            acc[a][b][j].x += tu.x;
            acc[a][b][j].y += tu.y;
            if (a == 0 && j == 0 && dual) {
              const float2 tu2 = cmul(tf2[b], u);
              acc2[b].x += tu2.x;
              acc2[b].y += tu2.y;
            }
          }
      if (c + 1 >= cend) {
        // item done: scatter to the parent grid and reset
        const int64_t obase = ((int64_t)p * F + f) * (int64_t)ntp * npp;
#pragma unroll
        for (int b = 0; b < CT; ++b) {
          const int col = k0 + 16 * b + c16;
          if (col < half) {
#pragma unroll
            for (int a = 0; a < ET; ++a)
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                const int e = e0 + 16 * a + 4 * g + j;
                if (e < n2p) {
                  int row = e, cc = col;
                  if (e >= ntp) {
                    const int i2 = P.mirror - e;  // in [0, ntp]
                    row = i2 < ntp ? i2 : -1;
                    cc = col + half;
                  }
                  if (row >= 0) {
                    const int64_t o = obase + (int64_t)row * npp + cc;
                    if (ROOT)
                      P.out_root[o] = make_double2((double)acc[a][b][j].x, (double)acc[a][b][j].y);
                    else
                      P.out[o] = make_float2((float)acc[a][b][j].x, (float)acc[a][b][j].y);
                  }
                  if (a == 0 && j == 0 && dual) {
                    const int64_t o2 = obase + (int64_t)i2_0 * npp + col + half;
                    if (ROOT)
                      P.out_root[o2] = make_double2(P.dual_sign * (double)acc2[b].x,
                                                    P.dual_sign * (double)acc2[b].y);
                    else
                      P.out[o2] = make_float2(P.dual_sign * (float)acc2[b].x,
                                              P.dual_sign * (float)acc2[b].y);
                  }
                }
              }
          }
        }
#pragma unroll
        for (int a = 0; a < ET; ++a)
#pragma unroll
          for (int b = 0; b < CT; ++b)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              acc[a][b][j].x = 0;
              acc[a][b][j].y = 0;
            }
#pragma unroll
        for (int b = 0; b < CT; ++b) {
          acc2[b].x = 0;
          acc2[b].y = 0;
        }
      }
      {
        int nit = it, np_ = p, nf = f, nc = c, ncend = cend;
        advance(nit, np_, nf, nc, ncend);
        live = nit < P.nitems;
        it = nit;
        p = np_;
        f = nf;
        c = nc;
        cend = ncend;
      }
      buf ^= 1;
    }
  }
}

// Extended translation table: Text[o][e][k], e < 2 ntp, k < npp/2, holds the
// factor of the parent sample fed by great-circle row e of column pair k (zero
// where none); row 2 ntp holds the second sample fed by e = 0 (see above).
// kmajor: the same factors stored Text[o][k][e] (e contiguous: the LDS FFT
// theta kernel reads one great circle per wave).
__global__ __launch_bounds__(256) void trans_ext_kernel(float2* __restrict__ Text, int ntp, int npp,
                                                        int mirror, double theta0, double hx,
                                                        double hy, double hz, double k_rev,
                                                        int kmajor) {
  const int half = npp / 2, n2p = 2 * ntp;
  const int64_t per_o = (int64_t)(n2p + 1) * half;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 8 * per_o;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int o = (int)(t / per_o);
    const int64_t d = t - (int64_t)o * per_o;
    int e, k;
    if (kmajor) {
      k = (int)(d / (n2p + 1));
      e = (int)(d - (int64_t)k * (n2p + 1));
    } else {
      e = (int)(d / half);
      k = (int)(d - (int64_t)e * half);
    }
    int row = -1, col = k;
    if (e < n2p && e < ntp) {
      row = e;
    } else {
      const int ee = (e == n2p) ? 0 : e;
      int i2 = (mirror - ee) % n2p;
      if (i2 < 0) i2 += n2p;
      if (i2 < ntp) {
        row = i2;
        col = k + half;
      }
    }
    float2 v = make_float2(0.f, 0.f);
    if (row >= 0) {
      const double th = theta0 + WF_PI / ntp * row;
      const double ph = WF_TWO_PI / npp * col;
      double st, ct, sp, cp;
      sincos(th, &st, &ct);
      sincos(ph, &sp, &cp);
      const double sx = (o & 1) ? 1.0 : -1.0;
      const double sy = (o & 2) ? -1.0 : 1.0;
      const double sz = (o & 4) ? -1.0 : 1.0;
      v = cis_rev(k_rev * (sx * hx * ct * cp + sy * hy * ct * sp + sz * hz * st));
    }
    Text[t] = v;
  }
}

// Per-child extended translation table (children at arbitrary offsets from
// the parent centre, e.g. the level-r nodes feeding the output grid
// directly): T[c][e][k] = exp(+i k0 s(e, k) . d_c), same rows as above.
__global__ __launch_bounds__(256) void trans_child_kernel(float2* __restrict__ T, int ntp, int npp,
                                                          int mirror, double theta0,
                                                          const double* __restrict__ cx,
                                                          const double* __restrict__ cy,
                                                          const double* __restrict__ cz, int nch,
                                                          double k_rev, int kmajor) {
  const int half = npp / 2, n2p = 2 * ntp;
  const int64_t per_c = (int64_t)(n2p + 1) * half;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < per_c * nch;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t / per_c);
    const int64_t d = t - (int64_t)c * per_c;
    int e, k;
    if (kmajor) {
      k = (int)(d / (n2p + 1));
      e = (int)(d - (int64_t)k * (n2p + 1));
    } else {
      e = (int)(d / half);
      k = (int)(d - (int64_t)e * half);
    }
    int row = -1, col = k;
    if (e < n2p && e < ntp) {
      row = e;
    } else {
      const int ee = (e == n2p) ? 0 : e;
      int i2 = (mirror - ee) % n2p;
      if (i2 < 0) i2 += n2p;
      if (i2 < ntp) {
        row = i2;
        col = k + half;
      }
    }
    float2 v = make_float2(0.f, 0.f);
    if (row >= 0) {
      const double th = theta0 + WF_PI / ntp * row;
      const double ph = WF_TWO_PI / npp * col;
      double st, ct, sp, cp;
      sincos(th, &st, &ct);
      sincos(ph, &sp, &cp);
      v = cis_rev(k_rev * (cx[c] * ct * cp + cy[c] * ct * sp + cz[c] * st));
    }
    T[t] = v;
  }
}

// Offsets d_c = c_node - c_root of the level-r nodes from their keys (r
// octant digits, level 1 most significant): the digit o of level l + 1 moves
// the centre by (sx hx_l, sy hy_l, sz hz_l), sx = o&1 ? 1 : -1, sy = o&2 ? -1
// : 1, sz = o&4 ? -1 : 1 (the octant code of the split planes).
struct LevelOffsets {
  double hx[16], hy[16], hz[16];
};
__global__ __launch_bounds__(256) void child_offsets_kernel(const uint32_t* __restrict__ keys,
                                                            int n, int r, LevelOffsets H,
                                                            double* __restrict__ cx,
                                                            double* __restrict__ cy,
                                                            double* __restrict__ cz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t key = keys[i];
  double x = 0, y = 0, z = 0;
  for (int l = 0; l < r; ++l) {
    const uint32_t o = (key >> (3 * (r - 1 - l))) & 7u;
    x += (o & 1) ? H.hx[l] : -H.hx[l];
    y += (o & 2) ? -H.hy[l] : H.hy[l];
    z += (o & 4) ? -H.hz[l] : H.hz[l];
  }
  cx[i] = x;
  cy[i] = y;
  cz[i] = z;
}


// blocks_per_cu: resident blocks aimed at per CU (sets the item stride)
template <int ET, int CT, int NKS, int D, bool ROOT, int MINB>
hipError_t launch_theta_t(const wf::ThetaMfmaArgs& a, int blocks_per_cu, hipStream_t st) {
  constexpr int BE = 4 * ET * 16, BC = CT * 16;
  ThetaK P;
  P.GT = reinterpret_cast<const uint4*>(a.GT);
  P.F = a.F;
  P.ntc = a.ntc;
  P.npp = a.npp;
  P.ntp = a.ntp;
  P.n2c8 = wf::gc_row_len(a.ntc);
  const int half = a.npp / 2;
  P.neb = (2 * a.ntp + BE - 1) / BE;
  P.ncb = (half + BC - 1) / BC;
  const int64_t nitems = (int64_t)a.nparents * a.F;
  if (nitems > 0x7fffffffLL) return hipErrorInvalidValue;
  P.nitems = (int)nitems;
  int64_t ng = std::max<int64_t>(1, 256LL * blocks_per_cu / ((int64_t)P.neb * P.ncb));
  ng = std::min<int64_t>(ng, nitems);
  P.ng = (int)ng;
  P.ngF = P.ng / a.F;
  P.ngR = P.ng % a.F;
  P.used = (int64_t)P.neb * P.ncb * ng;
  const int64_t nb = ((P.used + 7) / 8) * 8;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  P.kpad = a.kpad;
  P.mrows = wf::interp_rows_pad(2 * a.ntp);
  P.mirror = a.mirror;
  P.cb = a.cb;
  P.ce = a.ce;
  P.child_key = a.child_key;
  P.Text = a.Text;
  P.text_per_child = a.text_per_child;
  P.Mhi = reinterpret_cast<const bf16x8*>(a.Mhi);
  P.Mlo = reinterpret_cast<const bf16x8*>(a.Mlo);
  P.out = a.out;
  P.out_root = a.out_root;
  P.dual_sign = a.flip ? -1.f : 1.f;
  hipLaunchKernelGGL((theta_mfma_kernel<ET, CT, NKS, D, ROOT, MINB>), dim3((unsigned)nb), dim3(256), 0,
                     st, P);
  return hipGetLastError();
}

}  // namespace

namespace wf {

int interp_rows_pad(int rows) { return ((rows + 63) / 64) * 64; }
int interp_kpad(int k) { return ((k + 31) / 32) * 32; }

namespace {
uint16_t bf16_bits(float f) {  // round to nearest even (finite inputs)
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
float bf16_value(uint16_t b) {
  const uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}
// Periodic Dirichlet kernel of the band of n samples (Nyquist bin halved for even n).
double dirichlet(int n, double d) {
  d = std::remainder(d, 2.0 * WF_PI);  // (-pi, pi]
  const double s = std::sin(0.5 * d);
  if (std::fabs(s) < 1e-14) return 1.0;
  if (n % 2) return std::sin(0.5 * n * d) / (n * s);
  return std::sin(0.5 * n * d) * std::cos(0.5 * d) / (n * s);
}
}  // namespace

void build_interp_matrix(int nc, int np, double shift, std::vector<uint16_t>& hi,
                         std::vector<uint16_t>& lo, int& rows_pad, int& kpad) {
  rows_pad = interp_rows_pad(np);
  kpad = interp_kpad(nc);
  hi.assign((size_t)rows_pad * kpad, 0);
  lo.assign((size_t)rows_pad * kpad, 0);
  for (int q = 0; q < np; ++q)
    for (int j = 0; j < nc; ++j) {
      const double v = dirichlet(nc, WF_TWO_PI * q / np + shift - WF_TWO_PI * j / nc);
      const uint16_t h = bf16_bits((float)v);
      hi[(size_t)q * kpad + j] = h;
      lo[(size_t)q * kpad + j] = bf16_bits((float)(v - (double)bf16_value(h)));
    }
}

void build_theta_matrix_gc(int ntc, int n2p, double shift, std::vector<uint16_t>& hi,
                           std::vector<uint16_t>& lo, int& rows_pad, int& kpad, bool flip) {
  const int nc = 2 * ntc, h8 = gc_half_len(ntc);
  rows_pad = interp_rows_pad(n2p);
  kpad = interp_kpad(2 * h8);
  hi.assign((size_t)rows_pad * kpad, 0);
  lo.assign((size_t)rows_pad * kpad, 0);
  for (int q = 0; q < n2p; ++q)
    for (int j = 0; j < 2 * h8; ++j) {
      const int t = j < h8 ? j : j - h8;  // theta of the stored sample
      if (t >= ntc) continue;
      const int pos = j < h8 ? t : nc - 1 - t;  // its great-circle position
      double v = dirichlet(nc, WF_TWO_PI * q / n2p + shift - WF_TWO_PI * pos / nc);
      if (flip && ((j >= h8) != (q >= n2p / 2))) v = -v;
      const uint16_t hb = bf16_bits((float)v);
      hi[(size_t)q * kpad + j] = hb;
      lo[(size_t)q * kpad + j] = bf16_bits((float)(v - (double)bf16_value(hb)));
    }
}

template <int NKS, int RT>
hipError_t launch_phi_tile(const float2* X, float2* GT, int64_t planes, int ntc, int npc, int npp,
                           const uint16_t* Mhi, const uint16_t* Mlo, int kpad, hipStream_t st) {
  constexpr int QT = 2, D = 1, BPC = 2;
  const int nqb = (npp + 64 * QT - 1) / (64 * QT);
  const int64_t nrt = (planes * gc_half_len(ntc) + RT - 1) / RT;
  int64_t ng = std::max<int64_t>(1, 256LL * BPC / nqb);
  ng = std::min<int64_t>(ng, nrt);
  const int64_t used = (int64_t)nqb * ng;
  const int64_t nb = ((used + 7) / 8) * 8;
  hipLaunchKernelGGL((phi_tile_kernel<NKS, QT, D, RT>), dim3((unsigned)nb), dim3(256), 0, st, X,
                     reinterpret_cast<bf16x4*>(GT), (int)planes, ntc, npc, npp,
                     reinterpret_cast<const bf16x8*>(Mhi), reinterpret_cast<const bf16x8*>(Mlo),
                     kpad, interp_rows_pad(npp), nqb, (int)ng, used);
  return hipGetLastError();
}

hipError_t launch_phi_mfma(const float2* X, float2* GT, int64_t rows, int ntc, int npc, int npp,
                           const uint16_t* Mhi, const uint16_t* Mlo, int kpad, hipStream_t st,
                           int row_tile) {
  if (rows <= 0) return hipSuccess;
  if (ntc < 1 || npc < 2 || (npc & 1) || npc > kPhiMfmaMaxK || npp < 2 || (npp & 1) ||
      kpad < npc || (kpad & 31) || rows % ntc)
    return hipErrorInvalidValue;
  const int64_t planes = rows / ntc;
  if (planes * gc_half_len(ntc) > 0x7fffffffLL - 64 ||
      (int64_t)(npp / 2) * gc_row_len(ntc) >= (1LL << 31) ||
      (int64_t)interp_rows_pad(npp) * kpad >= (1LL << 31))
    return hipErrorInvalidValue;
  if (row_tile != 0 && row_tile != 16 && row_tile != 32) return hipErrorInvalidValue;
  const int nks = (npc + 31) / 32;
  if (nks <= 2) return launch_phi_tile<2, 32>(X, GT, planes, ntc, npc, npp, Mhi, Mlo, kpad, st);
  if (nks == 3) return launch_phi_tile<3, 32>(X, GT, planes, ntc, npc, npp, Mhi, Mlo, kpad, st);
  if (nks == 4) return launch_phi_tile<4, 32>(X, GT, planes, ntc, npc, npp, Mhi, Mlo, kpad, st);
  if (nks == 5) return launch_phi_tile<5, 32>(X, GT, planes, ntc, npc, npp, Mhi, Mlo, kpad, st);
  // long rows: 16-row tiles on request (half the LDS)
  if (nks <= 7) {
    if (row_tile == 16)
      return launch_phi_tile<7, 16>(X, GT, planes, ntc, npc, npp, Mhi, Mlo, kpad, st);
    return launch_phi_tile<7, 32>(X, GT, planes, ntc, npc, npp, Mhi, Mlo, kpad, st);
  }
  if (row_tile == 16)
    return launch_phi_tile<9, 16>(X, GT, planes, ntc, npc, npp, Mhi, Mlo, kpad, st);
  return launch_phi_tile<9, 32>(X, GT, planes, ntc, npc, npp, Mhi, Mlo, kpad, st);
}

hipError_t launch_theta_mfma(const ThetaMfmaArgs& a, hipStream_t st) {
  if (a.nparents <= 0) return hipSuccess;
  if (a.ntc < 1 || a.ntp < 1 || a.npp < 2 || (a.npp & 1) || a.kpad < gc_row_len(a.ntc) ||
      (a.kpad & 31) || a.F < 1 || gc_row_len(a.ntc) > kThetaMfmaMaxK)
    return hipErrorInvalidValue;
  if (a.mirror != 2 * a.ntp - 1 && a.mirror != 2 * a.ntp) return hipErrorInvalidValue;
  if ((a.out_root != nullptr) == (a.out != nullptr)) return hipErrorInvalidValue;
  if ((int64_t)interp_rows_pad(2 * a.ntp) * a.kpad >= (1LL << 31) ||
      (int64_t)(a.npp / 2) * gc_row_len(a.ntc) >= (1LL << 31) ||
      (int64_t)(2 * a.ntp + 1) * 8 * (a.npp / 2) >= (1LL << 31))
    return hipErrorInvalidValue;
  if (a.tile < 0 || a.tile > 3) return hipErrorInvalidValue;
  // tile by K steps (32 great-circle samples each), measured on the sphere
  // and fiber level shapes; 6-9 K steps also take an explicit tile (rows x
  // column pairs): 1 = 128 x 32 (the default), 2 = 128 x 16, 3 = 64 x 16
  // (the two smaller ones leave LDS for two blocks per CU)
  const int nks = (gc_row_len(a.ntc) + 31) / 32;
  if (a.out_root != nullptr) {
    if (nks <= 4) return launch_theta_t<2, 1, 4, 2, true, 1>(a, 2, st);
    if (nks <= 9) {
      if (a.tile == 2) return launch_theta_t<2, 1, 9, 1, true, 1>(a, 2, st);
      if (a.tile == 3) return launch_theta_t<1, 1, 9, 1, true, 2>(a, 2, st);
      return launch_theta_t<2, 2, 9, 1, true, 1>(a, 2, st);
    }
    return launch_theta_t<1, 1, 16, 1, true, 1>(a, 1, st);
  }
  if (nks <= 2) return launch_theta_t<2, 1, 2, 2, false, 2>(a, 2, st);
  if (nks == 3) return launch_theta_t<2, 1, 3, 2, false, 2>(a, 2, st);
  if (nks == 4) return launch_theta_t<2, 1, 4, 2, false, 2>(a, 2, st);
  if (nks == 5) return launch_theta_t<1, 1, 5, 2, false, 2>(a, 4, st);
  if (nks <= 9) {
    if (a.tile == 2) return launch_theta_t<2, 1, 9, 1, false, 1>(a, 2, st);
    if (a.tile == 3) return launch_theta_t<1, 1, 9, 1, false, 2>(a, 2, st);
    return launch_theta_t<2, 2, 9, 1, false, 1>(a, 2, st);
  }
  return launch_theta_t<1, 1, 16, 1, false, 1>(a, 1, st);
}

hipError_t launch_trans_ext(float2* Text, int ntp, int npp, int mirror, double theta0,
                            const double h[3], double k0, hipStream_t st, int kmajor) {
  const int64_t tot = 8LL * (2 * ntp + 1) * (npp / 2);
  int64_t blocks = (tot + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) return hipSuccess;
  hipLaunchKernelGGL(trans_ext_kernel, dim3((unsigned)blocks), dim3(256), 0, st, Text, ntp, npp,
                     mirror, theta0, h[0], h[1], h[2], k0 / WF_TWO_PI, kmajor);
  return hipGetLastError();
}

hipError_t launch_trans_child(float2* T, int ntp, int npp, int mirror, double theta0,
                              const double* cx, const double* cy, const double* cz, int nch,
                              double k0, hipStream_t st, int kmajor) {
  const int64_t tot = (int64_t)nch * (2 * ntp + 1) * (npp / 2);
  int64_t blocks = (tot + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) return hipSuccess;
  hipLaunchKernelGGL(trans_child_kernel, dim3((unsigned)blocks), dim3(256), 0, st, T, ntp, npp,
                     mirror, theta0, cx, cy, cz, nch, k0 / WF_TWO_PI, kmajor);
  return hipGetLastError();
}

hipError_t launch_child_offsets(const uint32_t* keys, int n, int r, const double* hx,
                                const double* hy, const double* hz, double* cx, double* cy,
                                double* cz, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  if (r < 1 || r > 10) return hipErrorInvalidValue;
  LevelOffsets H;
  for (int l = 0; l < 16; ++l) {
    H.hx[l] = l < r ? hx[l] : 0.0;
    H.hy[l] = l < r ? hy[l] : 0.0;
    H.hz[l] = l < r ? hz[l] : 0.0;
  }
  hipLaunchKernelGGL(child_offsets_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     keys, n, r, H, cx, cy, cz);
  return hipGetLastError();
}

}  // namespace wf
