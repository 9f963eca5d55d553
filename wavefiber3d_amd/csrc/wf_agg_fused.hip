// wf_agg_fused.hip -- one level transition of the tree aggregation in a
// single kernel: phi interpolation, theta interpolation over the great
// circle, translation and child sum, with the phi-pass output kept in LDS.
//
// The reference resamples every (child, component) pattern with cuFFT plans
// (interpolation.h:11-92), translates it (fields.h:92-103) and adds it into
// the parent (fields.h:153-170), one thrust call or kernel per step.  Here
// resampling is the dense Dirichlet-kernel map of the MFMA transition
// (wf_agg_mfma.hip): for a child pattern X [ntc][npc],
//
//   G[theta][q]  = sum_j X[theta][j] Mphi[q][j]                   (q < npp)
//   U[e][k]      = sum_j' Mth[e][j'] G'[j'][k]                   (e < 2 ntp)
//   parent(e, k) += T_octant(e, k) U[e][k]
//
// with G' the column pair (k, k + npp/2) stored as one great-circle row
// (column k theta-ascending, then column k + npp/2 theta-ascending, each
// padded to gc_half_len(ntc); the theta matrix's columns are permuted to
// that order).  Both GEMMs run on v_mfma_f32_16x16x32_bf16 with split BF16
// x3 operands (hi*hi + hi*lo + lo*hi) and FP32 accumulation.
//
// Work item = (parent, field, pair tile of up to 64 column pairs).  A block
// of NW waves walks items with a grid stride; for every child of the item:
//   1. phi GEMM: units (16 child theta rows) x (16 parent phi columns of the
//      tile) spread over the waves; the A operand (child rows) is read
//      straight from HBM into registers, the B operand (phi matrix rows) from
//      L2; the result is split into BF16 hi/lo and written to LDS in the B
//      layout of the theta GEMM;
//   2. barrier;
//   3. theta GEMM: wave w owns e-tiles w, w + NW (the theta matrix
//      fragments stay in registers for the block's life); per 16-pair tile
//      of columns, the MFMA result is multiplied by the translation factor
//      of the child's octant and added into the parent accumulators;
//   4. barrier.
// Each parent sample is written once, after its last child.  The compulsory
// HBM traffic is what remains: every child pattern read once per pair tile
// (once when npp/2 <= 64) and every parent written once.
#include <stdint.h>

#include <algorithm>

#include "wf_common.h"
#include "wf_interp.h"

namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int kMaxK = 4;              // 16-pair column tiles per item
constexpr int kNcol = 16 * kMaxK;     // column pairs per item

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

struct FusedK {
  const float2* X;  // child patterns [child][F][ntc][npc]
  float2* out;      // parent patterns [parent][F][ntp][npp]
  int F, ntc, npc, ntp, npp;
  int half, h8, nth, net, npt;  // npp/2, gc_half_len(ntc), child theta tiles, e-tiles, pair tiles
  int nkp, nkt;                 // K steps of the phi and theta GEMMs in use
  int mirror;
  int64_t nitems, used;
  int64_t nparents_f;  // parents x fields
  const int32_t* cb;
  const int32_t* ce;
  const uint32_t* child_key;
  const float2* Text;  // [8][2 ntp + 1][npp/2]
  const bf16x8* Mphi_hi;
  const bf16x8* Mphi_lo;
  int kpad_phi, mrows_phi;
  const bf16x8* Mth_hi;
  const bf16x8* Mth_lo;
  int kpad_th, mrows_th;
};

// NKP / NKT: maximum K steps (32 samples) of the phi / theta GEMMs; MAXE:
// e-tiles per wave.  LDS: the theta GEMM's B operand for one child,
// [NKT][4 planes][kNcol columns][4 k groups] of 8 BF16 (planes: re hi,
// re lo, im hi, im lo), so an MFMA lane (column c16, k group g) reads 16
// consecutive bytes and a wave reads 1 KB contiguously.
template <int NKP, int NKT, int MAXE>
__global__ __launch_bounds__(MAXE == 1 ? 768 : 512) void level_fused_mfma_kernel(const FusedK P) {
  __shared__ bf16x8 Gs[NKT * 4 * kNcol * 4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int NW = blockDim.x >> 6;
  const int c16 = lane & 15, g = lane >> 4;
  // XCD-aware: blocks b = x (mod 8) run on XCD x and take a contiguous range
  const int64_t nb = gridDim.x, per = (nb + 7) / 8;
  const int64_t w = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (w >= P.used) return;  // whole block
  const int F = P.F, ntc = P.ntc, npc = P.npc, ntp = P.ntp, npp = P.npp;
  const int half = P.half, h8 = P.h8, n2p = 2 * ntp;
  // the padding of the LDS operand (columns beyond the pair tile, K beyond
  // the great-circle row) stays zero for the block's life
  {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
    for (int i = threadIdx.x; i < NKT * 4 * kNcol * 4; i += blockDim.x) Gs[i] = z;
  }
  // resident theta-matrix fragments: A lane (row e = 16 et + c16, k group g)
  bf16x8 ah[MAXE][NKT], al[MAXE][NKT];
#pragma unroll
  for (int a = 0; a < MAXE; ++a) {
    const int et = wave + a * NW;
    const int row = 16 * et + c16;
    const int rr = row < P.mrows_th ? row : 0;
    const bf16x8* ph = P.Mth_hi + (((int64_t)rr * P.kpad_th) >> 3);
    const bf16x8* pl = P.Mth_lo + (((int64_t)rr * P.kpad_th) >> 3);
#pragma unroll
    for (int ks = 0; ks < NKT; ++ks) {
      const int kq = ks < P.nkt ? ks * 4 + g : 0;
      ah[a][ks] = ph[kq];
      al[a][ks] = pl[kq];
    }
  }
  __syncthreads();
  const int nunits_max = P.nth * 2 * kMaxK;
  for (int64_t item = w; item < P.nitems; item += P.used) {
    const int pt = (int)(item % P.npt);
    const int64_t pf = item / P.npt;
    const int p = (int)(pf / F), f = (int)(pf - (int64_t)p * F);
    const int kbase = pt * kNcol;
    const int nk_here = min(kMaxK, (half - kbase + 15) >> 4);  // 16-pair tiles in use
    float2 acc[MAXE][kMaxK][4];
#pragma unroll
    for (int a = 0; a < MAXE; ++a)
#pragma unroll
      for (int b = 0; b < kMaxK; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[a][b][j] = make_float2(0.f, 0.f);
    const int c_begin = P.cb[p], c_end = P.ce[p];
    for (int c = c_begin; c < c_end; ++c) {
      const float2* Xc = P.X + ((int64_t)c * F + f) * ((int64_t)ntc * npc);
      // ---- phi GEMM into LDS
      for (int u = wave; u < nunits_max; u += NW) {
        const int t = u / (2 * kMaxK), n = u - t * (2 * kMaxK);
        const int nt16 = n % kMaxK;
        if (nt16 >= nk_here) continue;  // whole wave: no column of this tile in use
        const bool second = n >= kMaxK;
        const int kk = kbase + 16 * nt16 + c16;  // this lane's column pair
        const bool kok = kk < half;
        const int q = second ? kk + half : kk;  // its phi column
        const int qq = (kok && q < P.mrows_phi) ? q : 0;
        const int th = 16 * t + c16;  // A row of this lane
        const bool rok = th < ntc;
        const float2* xr = Xc + (int64_t)(rok ? th : 0) * npc;
        floatx4 dr = {0.f, 0.f, 0.f, 0.f}, di = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < NKP; ++ks) {
          if (ks < P.nkp) {
            const int j0 = 32 * ks + 8 * g;
            float vr[8], vi[8];
#pragma unroll
            for (int h = 0; h < 4; ++h) {
              const int jj = j0 + 2 * h;  // npc even: a pair never straddles the row end
              float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
              if (rok && jj < npc) v = *reinterpret_cast<const float4*>(xr + jj);
              vr[2 * h] = v.x;
              vi[2 * h] = v.y;
              vr[2 * h + 1] = v.z;
              vi[2 * h + 1] = v.w;
            }
            bf16x8 arh, arl, aih, ail;
            split8(vr, arh, arl);
            split8(vi, aih, ail);
            const int64_t mo = (((int64_t)qq * P.kpad_phi) >> 3) + ks * 4 + g;
            const bf16x8 bh = P.Mphi_hi[mo], bl = P.Mphi_lo[mo];
            dr = mfma3(arh, arl, bh, bl, dr);
            di = mfma3(aih, ail, bh, bl, di);
          }
        }
        // D lane: child theta rows 16 t + 4 g + j, column q; stored at
        // great-circle index (second ? h8 : 0) + theta of column pair kk
        const int th0 = 16 * t + 4 * g;
        if (kok && th0 < h8) {
          const int jg = (second ? h8 : 0) + th0;
          const int ks = jg >> 5, g8 = (jg >> 3) & 3, hoff = (jg >> 2) & 1;
          const int col = 16 * nt16 + c16;
          bf16x4 rh, rl, ih, il;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const __bf16 hr = (__bf16)dr[j], hi_ = (__bf16)di[j];
            rh[j] = hr;
            rl[j] = (__bf16)(dr[j] - (float)hr);
            ih[j] = hi_;
            il[j] = (__bf16)(di[j] - (float)hi_);
          }
          bf16x4* base = reinterpret_cast<bf16x4*>(Gs);
          const int unit = ((ks * 4) * kNcol + col) * 4 + g8;  // plane 0
          base[2 * unit + hoff] = rh;
          base[2 * (unit + 1 * kNcol * 4) + hoff] = rl;
          base[2 * (unit + 2 * kNcol * 4) + hoff] = ih;
          base[2 * (unit + 3 * kNcol * 4) + hoff] = il;
        }
      }
      __syncthreads();
      // ---- theta GEMM, translation, child sum
      const int oct = (int)(P.child_key[c] & 7u);
      const float2* To = P.Text + (int64_t)oct * (n2p + 1) * half;
#pragma unroll
      for (int b = 0; b < kMaxK; ++b) {
        if (b < nk_here) {
          const int kk = kbase + 16 * b + c16;
          const bool kok = kk < half;
          float2 tf[MAXE][4];
#pragma unroll
          for (int a = 0; a < MAXE; ++a)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int e = 16 * (wave + a * NW) + 4 * g + j;
              tf[a][j] = (kok && e < n2p) ? To[(int64_t)e * half + kk] : make_float2(0.f, 0.f);
            }
          floatx4 ur[MAXE], ui[MAXE];
#pragma unroll
          for (int a = 0; a < MAXE; ++a) {
            ur[a] = floatx4{0.f, 0.f, 0.f, 0.f};
            ui[a] = floatx4{0.f, 0.f, 0.f, 0.f};
          }
#pragma unroll
          for (int ks = 0; ks < NKT; ++ks) {
            if (ks < P.nkt) {
              const int col = 16 * b + c16;
              const bf16x8 brh = Gs[((ks * 4 + 0) * kNcol + col) * 4 + g];
              const bf16x8 brl = Gs[((ks * 4 + 1) * kNcol + col) * 4 + g];
              const bf16x8 bih = Gs[((ks * 4 + 2) * kNcol + col) * 4 + g];
              const bf16x8 bil = Gs[((ks * 4 + 3) * kNcol + col) * 4 + g];
#pragma unroll
              for (int a = 0; a < MAXE; ++a) {
                ur[a] = mfma3(ah[a][ks], al[a][ks], brh, brl, ur[a]);
                ui[a] = mfma3(ah[a][ks], al[a][ks], bih, bil, ui[a]);
              }
            }
          }
#pragma unroll
          for (int a = 0; a < MAXE; ++a)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const float2 tu = cmul(tf[a][j], make_float2(ur[a][j], ui[a][j]));
              acc[a][b][j].x += tu.x;
              acc[a][b][j].y += tu.y;
            }
        }
      }
      __syncthreads();  // the next child's phi pass overwrites the LDS operand
    }
    // ---- the parent samples of this item, written once
    float2* O = P.out + ((int64_t)p * F + f) * ((int64_t)ntp * npp);
#pragma unroll
    for (int a = 0; a < MAXE; ++a)
#pragma unroll
      for (int b = 0; b < kMaxK; ++b) {
        const int kk = kbase + 16 * b + c16;
        if (b < nk_here && kk < half) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int e = 16 * (wave + a * NW) + 4 * g + j;
            if (e < n2p) {
              int row = e, col = kk;
              if (e >= ntp) {
                row = P.mirror - e;
                col = kk + half;
              }
              if (row >= 0 && row < ntp) O[(int64_t)row * npp + col] = acc[a][b][j];
            }
          }
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Second form of the same transition, built for occupancy and few VALU
// instructions per MFMA: a block keeps one pair tile for its life, with that
// tile's phi-matrix fragments in LDS.  For child grids of up to RT 16-row
// tiles the child pattern is staged in LDS already split into BF16 hi / lo
// planes in the layout of the phi GEMM's A operand, so each sample is split
// once per (child, field) instead of once per output column tile; every
// thread loads its chunk (one row, 8 samples) of the next (child, field) into
// registers at the top of an iteration and splits and stores it after the
// theta GEMM, so the global loads of a child are issued a whole iteration
// before their use.  Larger child grids read the A operand from global
// memory and split it per unit (RT = 0).
constexpr int kChunk = 4;  // complex samples of a child row per staging chunk

template <int NKP, int NKT, int KMAX, int MAXE, int RT, int LB, int WPE>
__global__ __launch_bounds__(LB, WPE) void level_fused2_kernel(const FusedK P) {
  constexpr bool STAGE = RT > 0;
  constexpr int NCOL = 16 * KMAX;
  constexpr int R16 = 16 * (RT > 0 ? RT : 1);  // staged rows (zero beyond ntc)
  __shared__ bf16x8 Gs[NKT * 4 * NCOL * 4];
  __shared__ bf16x8 Ms[NKP * 2 * 2 * NCOL * 4];
  // LDS operands are [..][k group][column or row]: MFMA lanes with consecutive
  // c16 read consecutive 16 B (no bank conflicts).  Gs: [ks][plane: re hi,
  // re lo, im hi, im lo][k group][column]; Ms: [ks][hi, lo][k group][column
  // of both halves]; Xs: [ks][plane][k group][row]
  __shared__ bf16x8 Xs[STAGE ? NKP * 4 * R16 * 4 : 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int NW = blockDim.x >> 6;
  const int c16 = lane & 15, g = lane >> 4;
  const int64_t nb = gridDim.x, per = (nb + 7) / 8;
  const int64_t w = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (w >= P.used) return;  // whole block
  const int F = P.F, ntc = P.ntc, npc = P.npc, ntp = P.ntp, npp = P.npp;
  const int half = P.half, h8 = P.h8, n2p = 2 * ntp;
  const int npt = P.npt;
  const int pt = (int)(w % npt);
  const int64_t ub = P.used / npt;  // blocks per pair tile
  const int kbase = pt * NCOL;
  const int nk_here = min(KMAX, (half - kbase + 15) >> 4);
  {
    bf16x8 z;
#pragma unroll
    for (int j = 0; j < 8; ++j) z[j] = (__bf16)0.0f;
    for (int i = threadIdx.x; i < NKT * 4 * NCOL * 4; i += blockDim.x) Gs[i] = z;
    if (STAGE)
      for (int i = threadIdx.x; i < NKP * 4 * R16 * 4; i += blockDim.x) Xs[i] = z;
    // phi-matrix B fragments of this pair tile: unit (ks, hi/lo, k group, column)
    for (int i = threadIdx.x; i < NKP * 2 * 2 * NCOL * 4; i += blockDim.x) {
      const int col = i % (2 * NCOL), gg = (i / (2 * NCOL)) & 3, hl = (i / (8 * NCOL)) & 1,
                ks = i / (16 * NCOL);
      const int cc = col % NCOL;
      const int kk = kbase + cc;
      bf16x8 v = z;
      if (kk < half && ks < P.nkp) {
        const int q = (col < NCOL) ? kk : kk + half;
        const bf16x8* M = hl ? P.Mphi_lo : P.Mphi_hi;
        v = M[(((int64_t)q * P.kpad_phi) >> 3) + ks * 4 + gg];
      }
      Ms[i] = v;
    }
  }
  bf16x8 ah[MAXE][NKT], al[MAXE][NKT];
#pragma unroll
  for (int a = 0; a < MAXE; ++a) {
    const int et = wave + a * NW;
    const int row = 16 * et + c16;
    const int rr = row < P.mrows_th ? row : 0;
    const bf16x8* ph = P.Mth_hi + (((int64_t)rr * P.kpad_th) >> 3);
    const bf16x8* pl = P.Mth_lo + (((int64_t)rr * P.kpad_th) >> 3);
#pragma unroll
    for (int ks = 0; ks < NKT; ++ks) {
      const int kq = ks < P.nkt ? ks * 4 + g : 0;
      ah[a][ks] = ph[kq];
      al[a][ks] = pl[kq];
    }
  }
  // staging chunk of this thread: row cr, samples kChunk cq .. kChunk cq + 3
  const int nq = (npc + kChunk - 1) / kChunk;
  const bool has_chunk = STAGE && (int)threadIdx.x < ntc * nq;
  // consecutive threads take consecutive rows, so their LDS stores are
  // consecutive 16 B
  const int cq = (int)threadIdx.x / ntc, cr = (int)threadIdx.x - cq * ntc;
  float4 xr[kChunk / 2];
  // child c, field f: its plane starts at (c F + f) ntc npc
  // loads at clamped, in-range addresses (threads without a chunk read the
  // plane's first samples), values masked after: no divergent branch around
  // them, so they stay in flight until stash()
  auto prefetch_cf = [&](int c, int f) {
    if (STAGE) {
      const float2* row = P.X + ((int64_t)c * F + f) * ((int64_t)ntc * npc) +
                          (int64_t)(has_chunk ? cr : 0) * npc;
#pragma unroll
      for (int h = 0; h < kChunk / 2; ++h) {
        const int k = kChunk * cq + 2 * h;
        const bool ok = has_chunk && k < npc;
        const float4 v = *reinterpret_cast<const float4*>(row + (ok ? k : 0));
        xr[h] = ok ? v : make_float4(0.f, 0.f, 0.f, 0.f);
      }
    }
  };
  auto stash = [&]() {
    if (has_chunk) {
      // samples k = 4 cq + j: K step k / 32, k group (k / 8) % 4, half (k / 4) % 2
      bf16x4 rh, rl, ih, il;
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float v[4] = {xr[h].x, xr[h].y, xr[h].z, xr[h].w};  // re, im, re, im
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const __bf16 hr = (__bf16)v[2 * q], hi_ = (__bf16)v[2 * q + 1];
          rh[2 * h + q] = hr;
          rl[2 * h + q] = (__bf16)(v[2 * q] - (float)hr);
          ih[2 * h + q] = hi_;
          il[2 * h + q] = (__bf16)(v[2 * q + 1] - (float)hi_);
        }
      }
      const int base = (((cq >> 3) * 4) * 4 + ((cq >> 1) & 3)) * R16 + cr;
      bf16x4* X4 = reinterpret_cast<bf16x4*>(Xs);
      X4[2 * base + (cq & 1)] = rh;
      X4[2 * (base + 4 * R16) + (cq & 1)] = rl;
      X4[2 * (base + 2 * 4 * R16) + (cq & 1)] = ih;
      X4[2 * (base + 3 * 4 * R16) + (cq & 1)] = il;
    }
  };
  const int64_t nitems = (int64_t)P.nparents_f;
  const int nunits_max = P.nth * 2 * KMAX;
  int64_t pf = w / npt;
  __syncthreads();  // LDS zeroed before the first child is stored
  if (STAGE && pf < nitems) {
    const int p0 = (int)(pf / F), f0 = (int)(pf - (int64_t)p0 * F);
    if (P.cb[p0] < P.ce[p0]) {
      prefetch_cf(P.cb[p0], f0);
      stash();
    }
  }
  __syncthreads();
  for (; pf < nitems; pf += ub) {
    const int p = (int)(pf / F), f = (int)(pf - (int64_t)p * F);
    float2 acc[MAXE][KMAX][4];
#pragma unroll
    for (int a = 0; a < MAXE; ++a)
#pragma unroll
      for (int b = 0; b < KMAX; ++b)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[a][b][j] = make_float2(0.f, 0.f);
    const int c_begin = P.cb[p], c_end = P.ce[p];
    if (STAGE && c_begin == c_end && pf + ub < nitems) {  // no children: stage the next item's
      const int64_t pn = pf + ub;
      const int p2 = (int)(pn / F), f2 = (int)(pn - (int64_t)p2 * F);
      if (P.cb[p2] < P.ce[p2]) {
        prefetch_cf(P.cb[p2], f2);
        stash();
      }
      __syncthreads();
    }
    for (int c = c_begin; c < c_end; ++c) {
      // next (child, field) of this block: the next child of the item, else
      // the first child of the block's next item
      if (STAGE) {
        if (c + 1 < c_end) {
          prefetch_cf(c + 1, f);
        } else if (pf + ub < nitems) {
          const int64_t pn = pf + ub;
          const int p2 = (int)(pn / F), f2 = (int)(pn - (int64_t)p2 * F);
          if (P.cb[p2] < P.ce[p2]) prefetch_cf(P.cb[p2], f2);
        }
      }
      const int oct = (int)(P.child_key[c] & 7u);
      const float2* To = P.Text + (int64_t)oct * (n2p + 1) * half;
      const float2* Xc = P.X + ((int64_t)c * F + f) * ((int64_t)ntc * npc);
      // ---- phi GEMM into LDS
      // D lane (column c16, rows 4g..4g+3) of unit (t, nt16, half) -> its
      // split into BF16 hi/lo, written in the theta GEMM's B layout
      auto write_g = [&](const floatx4& dr, const floatx4& di, int t, int nt16, bool second) {
        const int kk = kbase + 16 * nt16 + c16;
        const int th0 = 16 * t + 4 * g;
        if (kk < half && th0 < h8) {
          const int jg = (second ? h8 : 0) + th0;
          const int ks = jg >> 5, g8 = (jg >> 3) & 3, hoff = (jg >> 2) & 1;
          const int col = 16 * nt16 + c16;
          bf16x4 rh, rl, ih, il;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const __bf16 hr = (__bf16)dr[j], hi_ = (__bf16)di[j];
            rh[j] = hr;
            rl[j] = (__bf16)(dr[j] - (float)hr);
            ih[j] = hi_;
            il[j] = (__bf16)(di[j] - (float)hi_);
          }
          bf16x4* base = reinterpret_cast<bf16x4*>(Gs);
          const int unit = ((ks * 4) * 4 + g8) * NCOL + col;
          base[2 * unit + hoff] = rh;
          base[2 * (unit + 1 * 4 * NCOL) + hoff] = rl;
          base[2 * (unit + 2 * 4 * NCOL) + hoff] = ih;
          base[2 * (unit + 3 * 4 * NCOL) + hoff] = il;
        }
      };
      if constexpr (STAGE) {
        // unit = (row tile, column tile of one half); A from the staged planes
        for (int u = wave; u < nunits_max; u += NW) {
          const int t = u / (2 * KMAX), n = u - t * (2 * KMAX);
          const int nt16 = n % KMAX;
          if (nt16 >= nk_here) continue;
          const int th = 16 * t + c16;
          floatx4 dr = {0.f, 0.f, 0.f, 0.f}, di = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
          for (int ks = 0; ks < NKP; ++ks) {
            if (ks < P.nkp) {
              const int ib = ((ks * 4) * 4 + g) * R16 + th;
              const bf16x8 arh = Xs[ib], arl = Xs[ib + 4 * R16], aih = Xs[ib + 2 * 4 * R16],
                           ail = Xs[ib + 3 * 4 * R16];
              const int colM = 16 * n + c16;
              const bf16x8 bh = Ms[((ks * 2 + 0) * 4 + g) * 2 * NCOL + colM];
              const bf16x8 bl = Ms[((ks * 2 + 1) * 4 + g) * 2 * NCOL + colM];
              dr = mfma3(arh, arl, bh, bl, dr);
              di = mfma3(aih, ail, bh, bl, di);
            }
          }
          write_g(dr, di, t, nt16, n >= KMAX);
        }
      } else {
        // unit = (row tile, column tile) for both halves: the A operand,
        // read from global memory, is split once for the two
        for (int u = wave; u < P.nth * KMAX; u += NW) {
          const int t = u / KMAX, nt16 = u - t * KMAX;
          if (nt16 >= nk_here) continue;
          const int th = 16 * t + c16;
          const bool rok = th < ntc;
          floatx4 dr0 = {0.f, 0.f, 0.f, 0.f}, di0 = dr0, dr1 = dr0, di1 = dr0;
#pragma unroll
          for (int ks = 0; ks < NKP; ++ks) {
            if (ks < P.nkp) {
              const int j0 = 32 * ks + 8 * g;
              float vr[8], vi[8];
#pragma unroll
              for (int h = 0; h < 4; ++h) {
                const int jj = j0 + 2 * h;
                float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
                if (rok && jj < npc) v = *reinterpret_cast<const float4*>(Xc + (int64_t)th * npc + jj);
                vr[2 * h] = v.x;
                vi[2 * h] = v.y;
                vr[2 * h + 1] = v.z;
                vi[2 * h + 1] = v.w;
              }
              bf16x8 arh, arl, aih, ail;
              split8(vr, arh, arl);
              split8(vi, aih, ail);
              const int c0 = 16 * nt16 + c16, c1 = c0 + NCOL;
              const bf16x8 bh0 = Ms[((ks * 2 + 0) * 4 + g) * 2 * NCOL + c0];
              const bf16x8 bl0 = Ms[((ks * 2 + 1) * 4 + g) * 2 * NCOL + c0];
              dr0 = mfma3(arh, arl, bh0, bl0, dr0);
              di0 = mfma3(aih, ail, bh0, bl0, di0);
              const bf16x8 bh1 = Ms[((ks * 2 + 0) * 4 + g) * 2 * NCOL + c1];
              const bf16x8 bl1 = Ms[((ks * 2 + 1) * 4 + g) * 2 * NCOL + c1];
              dr1 = mfma3(arh, arl, bh1, bl1, dr1);
              di1 = mfma3(aih, ail, bh1, bl1, di1);
            }
          }
          write_g(dr0, di0, t, nt16, false);
          write_g(dr1, di1, t, nt16, true);
        }
      }
      __syncthreads();
      // ---- theta GEMM, translation, child sum (waves past the last e-tile,
      // there to fill the block to a multiple of 4 waves, skip it)
#pragma unroll
      for (int b = 0; b < KMAX; ++b) {
        if (b < nk_here && 16 * wave < n2p) {
          const int kk = kbase + 16 * b + c16;
          float2 tf[MAXE][4];
#pragma unroll
          for (int a = 0; a < MAXE; ++a)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const int e = 16 * (wave + a * NW) + 4 * g + j;
              const bool ok = kk < half && e < n2p;  // clamped read, masked value
              const float2 v = To[ok ? (int64_t)e * half + kk : 0];
              tf[a][j] = ok ? v : make_float2(0.f, 0.f);
            }
          floatx4 ur[MAXE], ui[MAXE];
#pragma unroll
          for (int a = 0; a < MAXE; ++a) {
            ur[a] = floatx4{0.f, 0.f, 0.f, 0.f};
            ui[a] = floatx4{0.f, 0.f, 0.f, 0.f};
          }
#pragma unroll
          for (int ks = 0; ks < NKT; ++ks) {
            if (ks < P.nkt) {
              const int col = 16 * b + c16;
              const bf16x8 brh = Gs[((ks * 4 + 0) * 4 + g) * NCOL + col];
              const bf16x8 brl = Gs[((ks * 4 + 1) * 4 + g) * NCOL + col];
              const bf16x8 bih = Gs[((ks * 4 + 2) * 4 + g) * NCOL + col];
              const bf16x8 bil = Gs[((ks * 4 + 3) * 4 + g) * NCOL + col];
#pragma unroll
              for (int a = 0; a < MAXE; ++a) {
                ur[a] = mfma3(ah[a][ks], al[a][ks], brh, brl, ur[a]);
                ui[a] = mfma3(ah[a][ks], al[a][ks], bih, bil, ui[a]);
              }
            }
          }
#pragma unroll
          for (int a = 0; a < MAXE; ++a)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const float2 tu = cmul(tf[a][j], make_float2(ur[a][j], ui[a][j]));
              acc[a][b][j].x += tu.x;
              acc[a][b][j].y += tu.y;
            }
        }
      }
      if (STAGE) stash();  // every wave is past this child's phi GEMM (barrier above)
      __syncthreads();
    }
    float2* O = P.out + ((int64_t)p * F + f) * ((int64_t)ntp * npp);
#pragma unroll
    for (int a = 0; a < MAXE; ++a)
#pragma unroll
      for (int b = 0; b < KMAX; ++b) {
        const int kk = kbase + 16 * b + c16;
        if (b < nk_here && kk < half) {
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int e = 16 * (wave + a * NW) + 4 * g + j;
            if (e < n2p) {
              int row = e, col = kk;
              if (e >= ntp) {
                row = P.mirror - e;
                col = kk + half;
              }
              if (row >= 0 && row < ntp) O[(int64_t)row * npp + col] = acc[a][b][j];
            }
          }
        }
      }
  }
}

// LDS bytes of one block of the second form
constexpr int fused2_lds(int nkp, int nkt, int kmax, int rt) {
  return (nkt * 4 * 16 * kmax * 4 + nkp * 2 * 2 * 16 * kmax * 4 + (rt > 0 ? nkp * 4 * 16 * rt * 4 : 1)) *
         16;
}

template <int NKP, int NKT, int KMAX, int MAXE, int RT, int LB, int WPE>
hipError_t launch_fused2_t(const FusedK& P, int nw, hipStream_t st) {
  constexpr int lds = fused2_lds(NKP, NKT, KMAX, RT);
  static_assert(lds <= 160 * 1024, "LDS of the fused transition");
  // resident blocks per CU: by LDS, and by waves (WPE waves per SIMD)
  int bpc = std::max(1, (160 * 1024) / lds);
  bpc = std::min(bpc, std::max(1, (4 * WPE) / nw));
  FusedK Q = P;
  const int64_t pf_items = (int64_t)P.nparents_f;
  int64_t ub = std::max<int64_t>(1, std::min<int64_t>(256LL * bpc / P.npt, pf_items));
  Q.used = ub * P.npt;
  const int64_t nb = ((Q.used + 7) / 8) * 8;
  hipLaunchKernelGGL((level_fused2_kernel<NKP, NKT, KMAX, MAXE, RT, LB, WPE>), dim3((unsigned)nb),
                     dim3(64 * nw), 0, st, Q);
  return hipGetLastError();
}

// rt: staged row tiles (0: A operand from global memory).  Staged blocks
// have a multiple of 4 waves; two blocks of up to 8 waves in 80 KB of LDS
// each cap registers at 128 (4 waves per SIMD), one block of up to 12 at 168.
template <int NKP, int NKT, int LB>
hipError_t launch_fused2_kt(const FusedK& P, int kmax, int maxe, int rt, int nw, hipStream_t st) {
  if (maxe == 2) return launch_fused2_t<NKP, NKT, 4, 2, 0, 512, 1>(P, nw, st);
  if constexpr (fused2_lds(NKP, NKT, 3, 2) <= 80 * 1024 && LB == 512) {
    if (kmax == 3 && rt == 2) return launch_fused2_t<NKP, NKT, 3, 1, 2, 512, 4>(P, nw, st);
  }
  if constexpr (fused2_lds(NKP, NKT, 4, 2) <= 80 * 1024 && LB == 512) {
    if (kmax == 4 && rt == 2) return launch_fused2_t<NKP, NKT, 4, 1, 2, 512, 4>(P, nw, st);
  }
  if constexpr (fused2_lds(NKP, NKT, 4, 3) <= 160 * 1024) {
    if (rt > 0) {
      if (kmax == 3) return launch_fused2_t<NKP, NKT, 3, 1, 3, LB, 3>(P, nw, st);
      return launch_fused2_t<NKP, NKT, 4, 1, 3, LB, 3>(P, nw, st);
    }
  }
  if (kmax == 3) return launch_fused2_t<NKP, NKT, 3, 1, 0, LB, 1>(P, nw, st);
  return launch_fused2_t<NKP, NKT, 4, 1, 0, LB, 1>(P, nw, st);
}

template <int NKP, int LB>
hipError_t launch_fused2_kl(const FusedK& P, int nkt, int kmax, int maxe, int rt, int nw,
                            hipStream_t st) {
  if (nkt <= 2) return launch_fused2_kt<NKP, 2, LB>(P, kmax, maxe, rt, nw, st);
  if (nkt == 3) return launch_fused2_kt<NKP, 3, LB>(P, kmax, maxe, rt, nw, st);
  if (nkt == 4) return launch_fused2_kt<NKP, 4, LB>(P, kmax, maxe, rt, nw, st);
  return launch_fused2_kt<NKP, 5, LB>(P, kmax, maxe, rt, nw, st);
}

template <int NKP>
hipError_t launch_fused2_k(const FusedK& P, int nkt, int kmax, int maxe, int rt, int nw,
                           hipStream_t st) {
  if (nw <= 8) return launch_fused2_kl<NKP, 512>(P, nkt, kmax, maxe, rt, nw, st);
  return launch_fused2_kl<NKP, 768>(P, nkt, kmax, maxe, rt, nw, st);
}

template <int NKP, int NKT, int MAXE>
hipError_t launch_fused_t(const FusedK& P, int nw, hipStream_t st) {
  // resident blocks: LDS NKT * 16 KB per block, waves per SIMD by register use
  const int lds_kb = NKT * 16;
  int bpc = std::max(1, std::min(4, 160 / std::max(lds_kb, 1)));
  bpc = std::min(bpc, std::max(1, 16 / nw));
  FusedK Q = P;
  int64_t used = std::min<int64_t>(256LL * bpc, Q.nitems);
  Q.used = used;
  const int64_t nb = ((used + 7) / 8) * 8;
  hipLaunchKernelGGL((level_fused_mfma_kernel<NKP, NKT, MAXE>), dim3((unsigned)nb), dim3(64 * nw), 0, st,
                     Q);
  return hipGetLastError();
}

template <int NKP>
hipError_t launch_fused_k(const FusedK& P, int nkt, int maxe, int nw, hipStream_t st) {
  if (maxe == 1) {
    if (nkt <= 2) return launch_fused_t<NKP, 2, 1>(P, nw, st);
    if (nkt == 3) return launch_fused_t<NKP, 3, 1>(P, nw, st);
    if (nkt == 4) return launch_fused_t<NKP, 4, 1>(P, nw, st);
    return launch_fused_t<NKP, 5, 1>(P, nw, st);
  }
  if (nkt <= 2) return launch_fused_t<NKP, 2, 2>(P, nw, st);
  if (nkt == 3) return launch_fused_t<NKP, 3, 2>(P, nw, st);
  if (nkt == 4) return launch_fused_t<NKP, 4, 2>(P, nw, st);
  return launch_fused_t<NKP, 5, 2>(P, nw, st);
}

}  // namespace

namespace wf {

bool agg_fused_fits(int ntc, int npc, int ntp, int npp) {
  const int nkp = (npc + 31) / 32, nkt = (gc_row_len(ntc) + 31) / 32;
  const int net = (2 * ntp + 15) / 16;
  return ntc >= 1 && npc >= 2 && (npc % 2) == 0 && npp >= 2 && (npp % 2) == 0 && nkp <= 4 &&
         nkt <= 5 && net <= 16;
}

hipError_t launch_agg_fused_mfma(const AggFusedArgs& a, hipStream_t st) {
  if (a.nparents <= 0) return hipSuccess;
  if (!agg_fused_fits(a.ntc, a.npc, a.ntp, a.npp) || a.F < 1 || !a.cb || !a.ce || !a.child_key ||
      !a.Text)
    return hipErrorInvalidValue;
  if (a.mirror != 2 * a.ntp - 1) return hipErrorInvalidValue;  // internal levels only
  const int nkp = (a.npc + 31) / 32, nkt = (gc_row_len(a.ntc) + 31) / 32;
  if (a.kpad_phi < 32 * nkp || (a.kpad_phi & 31) || a.kpad_th < 32 * nkt || (a.kpad_th & 31) ||
      a.mrows_phi < a.npp || a.mrows_th < 2 * a.ntp)
    return hipErrorInvalidValue;
  FusedK P;
  P.X = a.X;
  P.out = a.out;
  P.F = a.F;
  P.ntc = a.ntc;
  P.npc = a.npc;
  P.ntp = a.ntp;
  P.npp = a.npp;
  P.half = a.npp / 2;
  P.h8 = gc_half_len(a.ntc);
  P.nth = (a.ntc + 15) / 16;
  P.net = (2 * a.ntp + 15) / 16;
  P.npt = (P.half + kNcol - 1) / kNcol;
  P.nkp = nkp;
  P.nkt = nkt;
  P.mirror = a.mirror;
  P.nitems = (int64_t)a.nparents * a.F * P.npt;
  P.used = 0;
  P.cb = a.cb;
  P.ce = a.ce;
  P.child_key = a.child_key;
  P.Text = a.Text;
  P.Mphi_hi = reinterpret_cast<const bf16x8*>(a.Mphi_hi);
  P.Mphi_lo = reinterpret_cast<const bf16x8*>(a.Mphi_lo);
  P.kpad_phi = a.kpad_phi;
  P.mrows_phi = a.mrows_phi;
  P.Mth_hi = reinterpret_cast<const bf16x8*>(a.Mth_hi);
  P.Mth_lo = reinterpret_cast<const bf16x8*>(a.Mth_lo);
  P.kpad_th = a.kpad_th;
  P.mrows_th = a.mrows_th;
  P.nparents_f = (int64_t)a.nparents * a.F;
  // one e-tile per wave up to 12 waves, else two
  const int maxe = P.net <= 12 ? 1 : 2;
  int nw = (P.net + maxe - 1) / maxe;
  if (a.version == 2) {
    // pair tiles of 48 columns when they cover the grid, else 64 (two
    // e-tiles per wave run the 64-column form only; a 16-wave block with one
    // e-tile per wave and 32-pair tiles measured 8 % slower on 72 x 128 ->
    // 125 x 240)
    const int kmax = (maxe == 1 && P.half <= 48) ? 3 : 4;
    P.npt = (P.half + 16 * kmax - 1) / (16 * kmax);
    P.nitems = P.nparents_f * P.npt;
    // stage the child in LDS when its rows fit 2 or 3 tiles and every
    // staging chunk has a thread
    const int nq = (a.npc + kChunk - 1) / kChunk;
    const int nw4 = (nw + 3) & ~3;  // waves of a staged block
    const int rt = (maxe == 1 && P.nth <= 3 && nw4 <= 16 && a.ntc * nq <= 64 * nw4)
                       ? std::max(2, P.nth)
                       : 0;
    if (rt > 0) nw = nw4;
    if (nkp <= 2) return launch_fused2_k<2>(P, nkt, kmax, maxe, rt, nw, st);
    if (nkp == 3) return launch_fused2_k<3>(P, nkt, kmax, maxe, rt, nw, st);
    return launch_fused2_k<4>(P, nkt, kmax, maxe, rt, nw, st);
  }
  if (nkp <= 2) return launch_fused_k<2>(P, nkt, maxe, nw, st);
  if (nkp == 3) return launch_fused_k<3>(P, nkt, maxe, nw, st);
  return launch_fused_k<4>(P, nkt, maxe, nw, st);
}

}  // namespace wf
