// wf_agg_mfma.hip -- one level transition of the tree aggregation as dense
// interpolation GEMMs on the MFMA units (rows A9, A11, A12).
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
// Both run on v_mfma_f32_16x16x32_bf16 with split-BF16 x3 operands (matrix
// split on the host, data split in registers), FP32 accumulation; the root
// level accumulates the child sum in FP64.  The matrix fragments come from
// L1/L2 (one small matrix per level).  The phi pass reads its data fragments
// straight from HBM in the MFMA lane layout and stores theta-contiguous
// (GT[cf][q][theta]); the theta pass stages each 32-sample step of the
// great-circle sequences once per block in LDS, pre-split, for all waves.
#include <stdint.h>

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

// phi pass.  Wave task = (MTW tiles of 16 complex rows) x (4 tiles of 16
// output columns).  A = data rows (re and im parts as two M tiles sharing one
// load), B = Mphi^T: lane (col q, k-group g) reads Mphi[q][k0 + 8g .. +7].
// Loads are branchless (clamped index + select); the next K step's data is
// loaded before the current step's MFMAs.
template <int MTW>
__global__ __launch_bounds__(256) void phi_mfma_kernel(
    const float2* __restrict__ X, float2* __restrict__ GT, int rows, int ntc, int npc, int npp,
    const bf16x8* __restrict__ Mhi, const bf16x8* __restrict__ Mlo, int kpad, int nng,
    int ntasks) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int task = blockIdx.x * 4 + wave;
  if (task >= ntasks) return;  // no barriers below
  const int ng = task % nng, mg = task / nng;
  const int c16 = lane & 15, g = lane >> 4;
  const int row0 = mg * 16 * MTW;
  floatx4 accr[MTW][4], acci[MTW][4];
#pragma unroll
  for (int m = 0; m < MTW; ++m)
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      accr[m][n] = floatx4{0.f, 0.f, 0.f, 0.f};
      acci[m][n] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
  const float2* xr[MTW];
  bool rok[MTW];
#pragma unroll
  for (int m = 0; m < MTW; ++m) {
    const int r = row0 + 16 * m + c16;
    rok[m] = r < rows;
    xr[m] = X + (int64_t)(rok[m] ? r : 0) * npc;
  }
  const bf16x8* mh[4];
  const bf16x8* ml[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) {
    const int q = ng * 64 + 16 * n + c16;  // < padded rows of Mphi
    mh[n] = Mhi + ((q * kpad) >> 3);
    ml[n] = Mlo + ((q * kpad) >> 3);
  }
  float2 nx[MTW][8];
  auto load = [&](int k0) {
    const int kb = k0 + 8 * g;
#pragma unroll
    for (int m = 0; m < MTW; ++m)
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const bool ok = rok[m] && (kb + j < npc);
        const float2 v = xr[m][ok ? kb + j : 0];
        nx[m][j] = ok ? v : make_float2(0.f, 0.f);
      }
  };
  load(0);
  for (int k0 = 0; k0 < npc; k0 += 32) {
    const int kb = k0 + 8 * g;
    bf16x8 arh[MTW], arl[MTW], aih[MTW], ail[MTW];
#pragma unroll
    for (int m = 0; m < MTW; ++m) {
      float vr[8], vi[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        vr[j] = nx[m][j].x;
        vi[j] = nx[m][j].y;
      }
      split8(vr, arh[m], arl[m]);
      split8(vi, aih[m], ail[m]);
    }
    if (k0 + 32 < npc) load(k0 + 32);  // in flight during the MFMAs below
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const bf16x8 bh = mh[n][kb >> 3], bl = ml[n][kb >> 3];
#pragma unroll
      for (int m = 0; m < MTW; ++m) {
        accr[m][n] = mfma3(arh[m], arl[m], bh, bl, accr[m][n]);
        acci[m][n] = mfma3(aih[m], ail[m], bh, bl, acci[m][n]);
      }
    }
  }
  // D lane: column q, rows 4g + j of each M tile
#pragma unroll
  for (int m = 0; m < MTW; ++m)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int r = row0 + 16 * m + 4 * g + j;
      if (r >= rows) continue;
      const int cf = r / ntc, th = r - cf * ntc;
      float2* o = GT + (int64_t)cf * npp * ntc + th;
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        const int q = ng * 64 + 16 * n + c16;
        if (q < npp) o[(int64_t)q * ntc] = make_float2(accr[m][n][j], acci[m][n][j]);
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

// theta pass + translation + child sum, block-tiled.  Block = 4 waves in
// WE x WC, wave tile ET x CT 16-tiles: BE = 16 WE ET great-circle rows e by
// BC = 16 WC CT column pairs k of one (parent p, field f); the child loop runs
// inside.  Per 32-K step the B operand (the child's great-circle sequences of
// the block's columns) is loaded once, coalesced, split into BF16 hi/lo and
// staged in LDS for all waves; the A operand (matrix rows) comes from L1/L2.
// Loads are branchless (clamped index + select).  The parent accumulates in
// registers in the D layout (lane: column k, rows 4g + j); row e of the
// extended grid feeds parent (e, k) when e < ntp and (mirror - e, k + npp/2)
// otherwise when that row is < ntp; with a sample at the pole (mirror =
// 2 ntp) row 0 also feeds (0, k + npp/2), kept in a second accumulator.
template <int ET, int CT, int WE, int WC, bool ROOT>
__global__ __launch_bounds__(256) void theta_mfma_kernel(
    const float2* __restrict__ GT, int F, int ntc, int npp, int ntp, const int32_t* __restrict__ cb,
    const int32_t* __restrict__ ce, const uint32_t* __restrict__ child_key,
    const float2* __restrict__ Text, const bf16x8* __restrict__ Mhi,
    const bf16x8* __restrict__ Mlo, int kpad, int mrows, int mirror, int neb, int ncb,
    int64_t total, float2* __restrict__ out, double2* __restrict__ out_root) {
  typedef typename AccT<ROOT>::type acc_t;
  constexpr int BE = WE * ET * 16, BC = WC * CT * 16;
  __shared__ bf16x8 Bs[BC][4][4];  // [column][k group][re hi, re lo, im hi, im lo]
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c16 = lane & 15, g = lane >> 4;
  const int we = wave / WC, wc = wave % WC;
  // XCD-aware mapping: blocks b = x (mod 8) run on XCD x and take a
  // contiguous range of work items, so the e/k blocks of one (p, f) -- which
  // read the same child data -- share an L2
  const int64_t nb = gridDim.x, per = (nb + 7) / 8;
  const int64_t w = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
  if (w >= total) return;  // whole block
  const int cbk = (int)(w % ncb);
  int64_t t = w / ncb;
  const int eb = (int)(t % neb);
  t /= neb;
  const int f = (int)(t % F), p = (int)(t / F);
  const int half = npp / 2, n2c = 2 * ntc, n2p = 2 * ntp;
  const int e0 = eb * BE + we * ET * 16;
  const int k0 = cbk * BC + wc * CT * 16;
  const int i2_0 = mirror % n2p;
  const bool dual = (i2_0 < ntp);
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
  // loader role (fixed per thread): column lc, k group lkg
  const bool ldr = threadIdx.x < BC * 4;
  const int lc = (threadIdx.x >> 2) % BC, lkg = threadIdx.x & 3;
  const int lk = cbk * BC + lc;
  const bool lkok = ldr && lk < half;
  const int off1 = (lkok ? lk : 0) * ntc;                   // column k, theta ascending
  const int off2 = (lkok ? lk + half : 0) * ntc + n2c - 1;  // column k + half, reversed
  const bf16x8* ah_p[ET];
  const bf16x8* al_p[ET];
#pragma unroll
  for (int a = 0; a < ET; ++a) {
    const int row = e0 + 16 * a + c16;
    const int rr = row < mrows ? row : 0;
    ah_p[a] = Mhi + ((rr * kpad) >> 3);
    al_p[a] = Mlo + ((rr * kpad) >> 3);
  }
  const int64_t slab = (int64_t)npp * ntc;
  const int c0 = cb ? cb[p] : p, c1 = cb ? ce[p] : p + 1;
  for (int c = c0; c < c1; ++c) {
    const float2* S = GT + ((int64_t)c * F + f) * slab;
    floatx4 ur[ET][CT], ui[ET][CT];
#pragma unroll
    for (int a = 0; a < ET; ++a)
#pragma unroll
      for (int b = 0; b < CT; ++b) {
        ur[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
        ui[a][b] = floatx4{0.f, 0.f, 0.f, 0.f};
      }
    for (int k0s = 0; k0s < n2c; k0s += 32) {
      __syncthreads();  // previous step's fragments consumed
      if (ldr) {
        const int jb = k0s + 8 * lkg;
        float vr[8], vi[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const int jp = jb + j;
          const bool ok = lkok && jp < n2c;
          const float2 v = S[ok ? (jp < ntc ? off1 + jp : off2 - jp) : 0];
          vr[j] = ok ? v.x : 0.f;
          vi[j] = ok ? v.y : 0.f;
        }
        bf16x8 a0, a1, a2, a3;
        split8(vr, a0, a1);
        split8(vi, a2, a3);
        Bs[lc][lkg][0] = a0;
        Bs[lc][lkg][1] = a1;
        Bs[lc][lkg][2] = a2;
        Bs[lc][lkg][3] = a3;
      }
      __syncthreads();
      const int kq = (k0s >> 3) + g;
      bf16x8 ah[ET], al[ET];
#pragma unroll
      for (int a = 0; a < ET; ++a) {
        ah[a] = ah_p[a][kq];
        al[a] = al_p[a][kq];
      }
#pragma unroll
      for (int b = 0; b < CT; ++b) {
        const int col = wc * CT * 16 + 16 * b + c16;
        const bf16x8 brh = Bs[col][g][0], brl = Bs[col][g][1], bih = Bs[col][g][2],
                     bil = Bs[col][g][3];
#pragma unroll
        for (int a = 0; a < ET; ++a) {
          ur[a][b] = mfma3(ah[a], al[a], brh, brl, ur[a][b]);
          ui[a][b] = mfma3(ah[a], al[a], bih, bil, ui[a][b]);
        }
      }
    }
    const int oct = child_key ? (int)(child_key[c] & 7u) : 0;
    const float2* To = Text ? Text + oct * (n2p + 1) * half : nullptr;
#pragma unroll
    for (int b = 0; b < CT; ++b) {
      const int col = k0 + 16 * b + c16;
      const bool cok = col < half;
#pragma unroll
      for (int a = 0; a < ET; ++a)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int e = e0 + 16 * a + 4 * g + j;
          const bool ok = cok && e < n2p;
          const float2 u = make_float2(ur[a][b][j], ui[a][b][j]);
          const float2 tu = To ? cmul(To[ok ? e * half + col : 0], u) : u;
          acc[a][b][j].x += ok ? tu.x : 0.f;
          acc[a][b][j].y += ok ? tu.y : 0.f;
          if (dual && ok && e == 0) {
            const float2 tu2 = To ? cmul(To[n2p * half + col], u) : u;
            acc2[b].x += tu2.x;
            acc2[b].y += tu2.y;
          }
        }
    }
  }
  // scatter to the parent grid
  const int64_t obase = ((int64_t)p * F + f) * (int64_t)ntp * npp;
#pragma unroll
  for (int b = 0; b < CT; ++b) {
    const int col = k0 + 16 * b + c16;
    if (col >= half) continue;
#pragma unroll
    for (int a = 0; a < ET; ++a)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = e0 + 16 * a + 4 * g + j;
        if (e >= n2p) continue;
        int row = e, cc = col;
        if (e >= ntp) {
          int i2 = mirror - e;  // mirror is 2 ntp - 1 or 2 ntp: i2 in [0, ntp]
          row = i2 < ntp ? i2 : -1;
          cc = col + half;
        }
        if (row >= 0) {
          const int64_t o = obase + (int64_t)row * npp + cc;
          if (ROOT)
            out_root[o] = make_double2((double)acc[a][b][j].x, (double)acc[a][b][j].y);
          else
            out[o] = make_float2((float)acc[a][b][j].x, (float)acc[a][b][j].y);
        }
        if (dual && e == 0) {
          const int64_t o2 = obase + (int64_t)i2_0 * npp + col + half;
          if (ROOT)
            out_root[o2] = make_double2((double)acc2[b].x, (double)acc2[b].y);
          else
            out[o2] = make_float2((float)acc2[b].x, (float)acc2[b].y);
        }
      }
  }
}

// Extended translation table: Text[o][e][k], e < 2 ntp, k < npp/2, holds the
// factor of the parent sample fed by great-circle row e of column pair k (zero
// where none); row 2 ntp holds the second sample fed by e = 0 (see above).
__global__ __launch_bounds__(256) void trans_ext_kernel(float2* __restrict__ Text, int ntp, int npp,
                                                        int mirror, double theta0, double hx,
                                                        double hy, double hz, double k_rev) {
  const int half = npp / 2, n2p = 2 * ntp;
  const int64_t per_o = (int64_t)(n2p + 1) * half;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 8 * per_o;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int o = (int)(t / per_o);
    const int64_t d = t - (int64_t)o * per_o;
    const int e = (int)(d / half), k = (int)(d - (int64_t)e * half);
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

int64_t blocks_for(int64_t tasks) { return (tasks + 3) / 4; }

template <int ET, int CT, int WE, int WC, bool ROOT>
hipError_t launch_theta_t(const wf::ThetaMfmaArgs& a, hipStream_t st) {
  constexpr int BE = WE * ET * 16, BC = WC * CT * 16;
  const int half = a.npp / 2;
  const int neb = (2 * a.ntp + BE - 1) / BE, ncb = (half + BC - 1) / BC;
  const int64_t total = (int64_t)a.nparents * a.F * neb * ncb;
  if (total <= 0) return hipSuccess;
  const int64_t nb = ((total + 7) / 8) * 8;
  if (nb > 0x7fffffffLL) return hipErrorInvalidValue;
  const int mrows = wf::interp_rows_pad(2 * a.ntp);
  hipLaunchKernelGGL((theta_mfma_kernel<ET, CT, WE, WC, ROOT>), dim3((unsigned)nb), dim3(256), 0, st,
                     a.GT, a.F, a.ntc, a.npp, a.ntp, a.cb, a.ce, a.child_key, a.Text,
                     reinterpret_cast<const bf16x8*>(a.Mhi), reinterpret_cast<const bf16x8*>(a.Mlo),
                     a.kpad, mrows, a.mirror, neb, ncb, total, a.out, a.out_root);
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

hipError_t launch_phi_mfma(const float2* X, float2* GT, int64_t rows, int ntc, int npc, int npp,
                           const uint16_t* Mhi, const uint16_t* Mlo, int kpad,
                           hipStream_t st) {
  if (rows <= 0) return hipSuccess;
  if (rows > 0x7fffffffLL - 64 || ntc < 1 || npc < 1 || npp < 1 || kpad < npc || (kpad & 31))
    return hipErrorInvalidValue;
  if ((int64_t)interp_rows_pad(npp) * kpad >= (1LL << 31)) return hipErrorInvalidValue;
  constexpr int MTW = 2;
  const int nng = (npp + 63) / 64;
  const int64_t nmg = (rows + 16 * MTW - 1) / (16 * MTW);
  const int64_t ntasks = nmg * nng;
  if (ntasks > 0x7fffffffLL - 4) return hipErrorInvalidValue;
  hipLaunchKernelGGL((phi_mfma_kernel<MTW>), dim3((unsigned)blocks_for(ntasks)), dim3(256), 0, st,
                     X, GT, (int)rows, ntc, npc, npp, reinterpret_cast<const bf16x8*>(Mhi),
                     reinterpret_cast<const bf16x8*>(Mlo), kpad, nng, (int)ntasks);
  return hipGetLastError();
}

hipError_t launch_theta_mfma(const ThetaMfmaArgs& a, hipStream_t st) {
  if (a.nparents <= 0) return hipSuccess;
  if (a.ntc < 1 || a.ntp < 1 || a.npp < 2 || (a.npp & 1) || a.kpad < 2 * a.ntc || (a.kpad & 31))
    return hipErrorInvalidValue;
  if (a.mirror != 2 * a.ntp - 1 && a.mirror != 2 * a.ntp) return hipErrorInvalidValue;
  if ((a.out_root != nullptr) == (a.out != nullptr)) return hipErrorInvalidValue;
  if ((int64_t)interp_rows_pad(2 * a.ntp) * a.kpad >= (1LL << 31) ||
      (int64_t)a.npp * a.ntc >= (1LL << 31))
    return hipErrorInvalidValue;
  // block tile by extended-grid height (measured on the fiber level shapes):
  // <= 96 rows: 64 x 16, <= 160: 64 x 32, <= 288: 96 x 32, else 64 x 64
  const int n2p = 2 * a.ntp;
  if (a.out_root != nullptr) return launch_theta_t<2, 2, 2, 2, true>(a, st);
  if (n2p <= 96) return launch_theta_t<1, 1, 4, 1, false>(a, st);
  if (n2p <= 160) return launch_theta_t<2, 1, 2, 2, false>(a, st);
  if (n2p <= 288) return launch_theta_t<3, 1, 2, 2, false>(a, st);
  return launch_theta_t<2, 2, 2, 2, false>(a, st);
}

hipError_t launch_trans_ext(float2* Text, int ntp, int npp, int mirror, double theta0,
                            const double h[3], double k0, hipStream_t st) {
  const int64_t tot = 8LL * (2 * ntp + 1) * (npp / 2);
  int64_t blocks = (tot + 255) / 256;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) return hipSuccess;
  hipLaunchKernelGGL(trans_ext_kernel, dim3((unsigned)blocks), dim3(256), 0, st, Text, ntp, npp,
                     mirror, theta0, h[0], h[1], h[2], k0 / WF_TWO_PI);
  return hipGetLastError();
}

}  // namespace wf
