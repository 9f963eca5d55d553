// wf_tree64.hip -- the PO tree in FP64, forward and reverse, for gradients.
//
// A loss of the log intensity weights every output direction by 1/I, so its
// gradient is only as accurate as the far field in the dimmest directions.
// With FP32 patterns the absolute error of every level is ~1e-7 of that
// level's largest values, which in directions 1e4 times dimmer than the peak
// is ~1e-3 relative: the FP32 tree's gradient stops at ~2e-4 of the FP64
// direct sum's however fine the sampling.  This file holds the same tree as
// the FP32 FFT path -- leaf sums (util/fields.h:8-38 summed per leaf as in
// treePO.cu:652-670), phi and great-circle theta interpolation with spectral
// pad/fold (interpolation.h:11-92), translation (fields.h:92-103), child sum
// (fields.h:153-170) -- and the exact adjoint of each step, with every
// pattern, table, twiddle and accumulator in FP64.  Phases are formed in FP64
// revolutions and reduced exactly before the FP64 sine/cosine (zcis_rev).
#include "wf_common.h"
#include "wf_tree64.h"

namespace {

// ------------------------------------------------------- FP64 LDS FFT --
__device__ inline double2 tw64(const double2* tw, int m, int inverse) {
  const double2 t = tw[m];
  return inverse ? make_double2(t.x, -t.y) : t;
}

template <int R>
__device__ inline void dft64(double2* v, const double2* tw, int step, int inverse) {
  if constexpr (R == 2) {
    const double2 a = v[0], b = v[1];
    v[0] = zadd(a, b);
    v[1] = zsub(a, b);
  } else if constexpr (R == 4) {
    const double2 a0 = zadd(v[0], v[2]), a1 = zsub(v[0], v[2]);
    const double2 b0 = zadd(v[1], v[3]), b1 = zsub(v[1], v[3]);
    const double2 b1r = inverse ? make_double2(-b1.y, b1.x) : make_double2(b1.y, -b1.x);
    v[0] = zadd(a0, b0);
    v[2] = zsub(a0, b0);
    v[1] = zadd(a1, b1r);
    v[3] = zsub(a1, b1r);
  } else {
    double2 o[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      double2 s = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) s = zadd(s, zmul(v[r], tw64(tw, ((q * r) % R) * step, inverse)));
      o[q] = s;
    }
#pragma unroll
    for (int q = 0; q < R; ++q) v[q] = o[q];
  }
}

// Stockham stage of radix R (the scheme of the FP32 LDS FFT, in FP64).
template <int R>
__device__ inline void stage64(const double2* __restrict__ src, double2* __restrict__ dst, int n,
                               int nlines, int lstride, int Ns, const double2* tw, int inverse) {
  const int m = n / R;
  const int total = nlines * m;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = t / m, j = t - line * m;
    const double2* s = src + line * lstride;
    double2* d = dst + line * lstride;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[j + r * m];
    const int k = j % Ns;
    if (Ns > 1) {
      const int tstep = n / (Ns * R);
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = zmul(v[r], tw64(tw, r * k * tstep, inverse));
    }
    dft64<R>(v, tw, m, inverse);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[base + r * Ns] = v[r];
  }
}

// Batched FFT of nlines lines; returns the buffer holding the result.
// Contains barriers: call from every thread of the block.
__device__ inline double2* lines64(double2* a, double2* b, int nlines, int lstride,
                                   const wf::Fft64Plan& p, int inverse) {
  int Ns = 1;
  for (int s = 0; s < p.nstages; ++s) {
    const int R = (int)((p.radix_packed >> (4 * s)) & 15u);
    switch (R) {
      case 2: stage64<2>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 3: stage64<3>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 4: stage64<4>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 5: stage64<5>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 7: stage64<7>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      case 11: stage64<11>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
      default: stage64<13>(a, b, p.n, nlines, lstride, Ns, p.tw, inverse); break;
    }
    __syncthreads();
    double2* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// The plan with its twiddle table copied into LDS at dst (p.n entries): the
// butterflies then read twiddles at LDS latency.  Call from every thread; a
// barrier must follow before the plan is used.
__device__ inline wf::Fft64Plan plan_in_lds(const wf::Fft64Plan& p, double2* dst) {
  for (int t = threadIdx.x; t < p.n; t += blockDim.x) dst[t] = p.tw[t];
  wf::Fft64Plan q = p;
  q.tw = dst;
  return q;
}

// Phase factors of the origin shift for the band of nc samples:
// ph[f + nc/2] = e^{i sign f shift}, f = -(nc/2) .. (nc-1)/2 (+ nc/2 for an even
// nc), into LDS; one FP64 sine/cosine per frequency instead of one per output
// sample and child.  Call from every thread; a barrier must follow.
__device__ inline void shift_phases64(double2* ph, int nc, double shift, double sign) {
  const int fmin = -(nc / 2), fmax = (nc - 1) / 2 + ((nc % 2 == 0) ? 1 : 0);
  for (int t = threadIdx.x; t <= fmax - fmin; t += blockDim.x)
    ph[t] = zcis_rev(sign * shift * (double)(fmin + t) / WF_TWO_PI);
}

// Spectral pad/fold of spectra X (length nc) onto np bins with the origin
// shift: Y[q] = scale sum_{f = q mod np, f in band(nc)} w_f e^{i f shift} X[f mod nc]
// (w_f = 1/2 on the two Nyquist halves of an even nc, else 1); ph: the
// factors of shift_phases64 (sign +1), or null for no shift.
__device__ inline void resample64(const double2* __restrict__ X, double2* __restrict__ Y,
                                  int nlines, int xs, int ys, int nc, int np,
                                  const double2* __restrict__ ph, double scale) {
  const int fmin = -(nc / 2), fmax = (nc - 1) / 2 + ((nc % 2 == 0) ? 1 : 0);
  const int total = nlines * np;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = t / np, q = t - line * np;
    const double2* x = X + line * xs;
    int f = q - ((q - fmin) / np) * np;
    if (f < fmin) f += np;
    double2 acc = make_double2(0.0, 0.0);
    for (; f <= fmax; f += np) {
      const int bin = f < 0 ? f + nc : f;
      double2 v = x[bin];
      if ((nc % 2 == 0) && (f == nc / 2 || f == -(nc / 2))) v = zscale(v, 0.5);
      if (ph) v = zmul(v, ph[f - fmin]);
      acc = zadd(acc, v);
    }
    Y[line * ys + q] = zscale(acc, scale);
  }
}

// Hermitian adjoint of resample64 (times scale); ph: shift_phases64 with
// sign -1, or null.
__device__ inline void resample64_adj(const double2* __restrict__ Y, double2* __restrict__ X,
                                      int nlines, int ys, int xs, int nc, int np,
                                      const double2* __restrict__ ph, double scale) {
  const int total = nlines * nc;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int line = t / nc, m = t - line * nc;
    const double2* y = Y + line * ys;
    double2 acc = make_double2(0.0, 0.0);
    const bool nyq = (nc % 2 == 0) && (m == nc / 2);
    const int nf = nyq ? 2 : 1;
    for (int h = 0; h < nf; ++h) {
      int f = (m <= (nc - 1) / 2) ? m : m - nc;
      if (nyq) f = h ? -(nc / 2) : nc / 2;
      int q = f % np;
      if (q < 0) q += np;
      double2 v = y[q];
      if (nyq) v = zscale(v, 0.5);
      if (ph) v = zmul(v, ph[f + nc / 2]);
      acc = zadd(acc, v);
    }
    X[line * xs + m] = zscale(acc, scale);
  }
}

__device__ inline void leaf_centre64(uint32_t key, int depth, const double lo[3],
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

// leaf-grid direction d (half-sample theta grid, the same as the FP32 leaf)
__device__ inline void leaf_dir64(int d, int nt, int np, double s[3]) {
  const int j = d / np, kk = d - j * np;
  const double th = -WF_PI / 2 + WF_PI * (j + 0.5) / nt;
  const double ph = WF_TWO_PI * kk / np;
  double st, ct, sp, cp;
  sincos(th, &st, &ct);
  sincos(ph, &sp, &cp);
  s[0] = ct * cp;
  s[1] = ct * sp;
  s[2] = st;
}

constexpr int kT = 256;    // threads per block
constexpr int kTile = 64;  // points (leaf) / directions (adjoint) staged per LDS tile
constexpr int kFB = 6;     // fields per leaf block (one solve)

// Leaf sums: out[leaf][f][d] = sum_{n in leaf} e^{+i k s_d.(x_n - c_leaf)} C[n][f].
__global__ __launch_bounds__(kT) void leaf64_kernel(wf::Leaf64Args a) {
  __shared__ double3 sr[kTile];
  __shared__ double2 sc[kTile][kFB];
  const int D = a.nt * a.np;
  const int tiles = (D + kT - 1) / kT;
  const int nfb = a.F / kFB;
  const int64_t bid = blockIdx.x;
  const int tile = (int)(bid % tiles);
  const int fb = (int)((bid / tiles) % nfb) * kFB;
  const int leaf = (int)(bid / ((int64_t)tiles * nfb));
  const int d = tile * kT + threadIdx.x;
  const int b = a.leaf_begin[leaf], e = a.leaf_end[leaf];
  double c[3];
  leaf_centre64(a.leaf_key[leaf], a.depth, a.box_lo, a.leaf_size, c);
  const double k_rev = a.k0 / WF_TWO_PI;
  double s[3] = {0.0, 0.0, 0.0};
  if (d < D) leaf_dir64(d, a.nt, a.np, s);
  double2 acc[kFB];
#pragma unroll
  for (int f = 0; f < kFB; ++f) acc[f] = make_double2(0.0, 0.0);
  for (int base = b; base < e; base += kTile) {
    const int cnt = min(kTile, e - base);
    __syncthreads();
    for (int t = threadIdx.x; t < cnt; t += kT) {
      const int i = base + t;
      sr[t] = make_double3(((double)a.px[i] - c[0]) * k_rev, ((double)a.py[i] - c[1]) * k_rev,
                           ((double)a.pz[i] - c[2]) * k_rev);
    }
    for (int t = threadIdx.x; t < cnt * kFB; t += kT) {
      const int p = t / kFB, f = t - p * kFB;
      sc[p][f] = a.C[(int64_t)(base + p) * a.F + fb + f];
    }
    __syncthreads();
#pragma unroll 2
    for (int p = 0; p < cnt; ++p) {  // two independent phase chains in flight
      const double3 r = sr[p];
      const double2 ph = zcis_rev(s[0] * r.x + s[1] * r.y + s[2] * r.z);
#pragma unroll
      for (int f = 0; f < kFB; ++f) acc[f] = zadd(acc[f], zmul(sc[p][f], ph));
    }
  }
  if (d < D) {
#pragma unroll
    for (int f = 0; f < kFB; ++f) a.out[((int64_t)leaf * a.F + fb + f) * D + d] = acc[f];
  }
}

// phi pass: [plane][ntc][npc] -> [plane][ntc][npp], rows_per_block rows per block.
__global__ __launch_bounds__(kT) void phi64_kernel(const double2* __restrict__ in,
                                                   double2* __restrict__ out, int ntc, int npc,
                                                   int npp, int rows_per_block, wf::Fft64Plan pc0,
                                                   wf::Fft64Plan pp0) {
  extern __shared__ double2 sm64[];
  const int lmax = npc > npp ? npc : npp;
  double2* a = sm64;
  double2* b = sm64 + rows_per_block * lmax;
  const wf::Fft64Plan pc = plan_in_lds(pc0, b + rows_per_block * lmax);
  const wf::Fft64Plan pp = plan_in_lds(pp0, b + rows_per_block * lmax + pc0.n);
  const int rblocks = (ntc + rows_per_block - 1) / rows_per_block;
  const int64_t plane = blockIdx.x / rblocks;
  const int row0 = (int)(blockIdx.x - plane * rblocks) * rows_per_block;
  const int nrows = min(rows_per_block, ntc - row0);
  const double2* src = in + (plane * ntc + row0) * (int64_t)npc;
  for (int t = threadIdx.x; t < nrows * npc; t += blockDim.x) {
    const int r = t / npc, k = t - r * npc;
    a[r * lmax + k] = src[(int64_t)r * npc + k];
  }
  __syncthreads();
  double2* X = lines64(a, b, nrows, lmax, pc, 0);
  double2* Y = (X == a) ? b : a;
  resample64(X, Y, nrows, lmax, lmax, npc, npp, nullptr, 1.0 / npc);
  __syncthreads();
  double2* Z = lines64(Y, X, nrows, lmax, pp, 1);
  double2* dst = out + (plane * ntc + row0) * (int64_t)npp;
  for (int t = threadIdx.x; t < nrows * npp; t += blockDim.x) {
    const int r = t / npp, k = t - r * npp;
    dst[(int64_t)r * npp + k] = Z[r * lmax + k];
  }
}

// theta pass + translation + child sum, one block per (parent, field, tile of
// W column pairs).  G: [child][F][ntc][npp]; out: [parent][F][ntp][npp] (the
// root: [F][ntp][npp]).  Parent slots in registers: thread t owns t + 256 m.
template <int MAXS>
__global__ __launch_bounds__(kT) void theta64_kernel(
    const double2* __restrict__ G, int F, int ntc, int npp, int ntp, int W,
    const int32_t* __restrict__ cb, const int32_t* __restrict__ ce,
    const uint32_t* __restrict__ child_key, const double2* __restrict__ T, double shift,
    int mirror_base, wf::Fft64Plan pc0, wf::Fft64Plan pp0, double2* __restrict__ out) {
  extern __shared__ double2 sm64[];
  const int n2c = 2 * ntc, n2p = 2 * ntp;
  const int lmax = n2c > n2p ? n2c : n2p;
  double2* a = sm64;
  double2* b = sm64 + W * lmax;
  double2* ph = sm64 + 2 * W * lmax;  // n2c + 1 shift factors, then the twiddles
  shift_phases64(ph, n2c, shift, 1.0);  // the child loop's first barrier follows
  const wf::Fft64Plan pc = plan_in_lds(pc0, ph + n2c + 1);
  const wf::Fft64Plan pp = plan_in_lds(pp0, ph + n2c + 1 + pc0.n);
  const int half = npp / 2;
  const int ntiles = (half + W - 1) / W;
  const int tile = (int)(blockIdx.x % ntiles);
  const int64_t pf = blockIdx.x / ntiles;
  const int f = (int)(pf % F);
  const int p = (int)(pf / F);
  const int k0 = tile * W;
  const int wcnt = min(W, half - k0);
  const int nslot = wcnt * n2p;
  double2 acc[MAXS];
#pragma unroll
  for (int m = 0; m < MAXS; ++m) acc[m] = make_double2(0.0, 0.0);
  const int c_lo = cb ? cb[p] : p, c_hi = ce ? ce[p] : p + 1;
  for (int c = c_lo; c < c_hi; ++c) {
    __syncthreads();
    const double2* g = G + ((int64_t)c * F + f) * ntc * (int64_t)npp;
    for (int t = threadIdx.x; t < ntc * wcnt; t += blockDim.x) {
      const int j = t / wcnt, w = t - j * wcnt;
      a[w * lmax + j] = g[(int64_t)j * npp + k0 + w];
      a[w * lmax + ntc + (ntc - 1 - j)] = g[(int64_t)j * npp + k0 + w + half];
    }
    __syncthreads();
    double2* X = lines64(a, b, wcnt, lmax, pc, 0);
    double2* Y = (X == a) ? b : a;
    resample64(X, Y, wcnt, lmax, lmax, n2c, n2p, ph, 1.0 / n2c);
    __syncthreads();
    double2* Z = lines64(Y, X, wcnt, lmax, pp, 1);
    const int oct = child_key ? (int)(child_key[c] & 7u) : 0;
    const double2* Tt = T ? T + (int64_t)oct * ntp * npp : nullptr;
#pragma unroll
    for (int m = 0; m < MAXS; ++m) {
      const int s = threadIdx.x + kT * m;
      if (s < nslot) {
        const int w = s / n2p, q = s - w * n2p;
        int i, col, e;
        if (q < ntp) {
          i = q;
          col = k0 + w;
          e = i;
        } else {
          i = q - ntp;
          col = k0 + w + half;
          e = (mirror_base - i) % n2p;
          if (e < 0) e += n2p;
        }
        double2 v = Z[w * lmax + e];
        if (Tt) v = zmul(v, Tt[(int64_t)i * npp + col]);
        acc[m] = zadd(acc[m], v);
      }
    }
  }
  double2* o = out + ((int64_t)p * F + f) * ntp * (int64_t)npp;
#pragma unroll
  for (int m = 0; m < MAXS; ++m) {
    const int s = threadIdx.x + kT * m;
    if (s < nslot) {
      const int w = s / n2p, q = s - w * n2p;
      const int i = q < ntp ? q : q - ntp;
      const int col = q < ntp ? k0 + w : k0 + w + half;
      o[(int64_t)i * npp + col] = acc[m];
    }
  }
}

// Translation table T[o][i][k] = exp(+i k0 s(theta_i, phi_k).dc_o) in FP64.
__global__ __launch_bounds__(kT) void trans64_kernel(double2* __restrict__ T, int ntp, int npp,
                                                     double theta0, double hx, double hy,
                                                     double hz, double k_rev) {
  const int64_t D = (int64_t)ntp * npp;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < 8 * D;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int o = (int)(t / D);
    const int64_t d = t - (int64_t)o * D;
    const int i = (int)(d / npp), k = (int)(d - (int64_t)i * npp);
    const double th = theta0 + WF_PI / ntp * i;
    const double ph = WF_TWO_PI / npp * k;
    double st, ct, sp, cp;
    sincos(th, &st, &ct);
    sincos(ph, &sp, &cp);
    const double sx = (o & 1) ? 1.0 : -1.0;
    const double sy = (o & 2) ? -1.0 : 1.0;
    const double sz = (o & 4) ? -1.0 : 1.0;
    T[t] = zcis_rev(k_rev * (sx * hx * ct * cp + sy * hy * ct * sp + sz * hz * st));
  }
}

// Adjoint of theta64: parent cogradient [parent][F][ntp][npp] -> phi-interpolated
// child cogradient [child][F][ntc][npp]; one block per (child, field, tile).
__global__ __launch_bounds__(kT) void theta_adj64_kernel(
    const double2* __restrict__ gP, int F, int ntc, int npp, int ntp, int W,
    const int32_t* __restrict__ parent_idx, const uint32_t* __restrict__ child_key,
    const double2* __restrict__ T, double shift, int mirror_base, wf::Fft64Plan pp0,
    wf::Fft64Plan pc0, double2* __restrict__ gG) {
  extern __shared__ double2 sm64[];
  const int n2c = 2 * ntc, n2p = 2 * ntp;
  const int lmax = n2c > n2p ? n2c : n2p;
  double2* a = sm64;
  double2* b = sm64 + W * lmax;
  double2* ph = sm64 + 2 * W * lmax;  // n2c + 1 shift factors (conjugate), then the twiddles
  shift_phases64(ph, n2c, shift, -1.0);  // the barrier after the load orders it
  const wf::Fft64Plan pp = plan_in_lds(pp0, ph + n2c + 1);
  const wf::Fft64Plan pc = plan_in_lds(pc0, ph + n2c + 1 + pp0.n);
  const int half = npp / 2;
  const int ntiles = (half + W - 1) / W;
  const int tile = (int)(blockIdx.x % ntiles);
  const int64_t cf = blockIdx.x / ntiles;
  const int f = (int)(cf % F);
  const int c = (int)(cf / F);
  const int k0 = tile * W;
  const int wcnt = min(W, half - k0);
  const int p = parent_idx ? parent_idx[c] : 0;
  const int oct = child_key ? (int)(child_key[c] & 7u) : 0;
  const double2* Tt = T ? T + (int64_t)oct * ntp * npp : nullptr;
  const int64_t pbase = ((int64_t)p * F + f) * ntp * (int64_t)npp;
  auto gp = [&](int i, int col) -> double2 {
    double2 v = gP[pbase + (int64_t)i * npp + col];
    if (Tt) {
      const double2 t = Tt[(int64_t)i * npp + col];
      v = zmul(make_double2(t.x, -t.y), v);
    }
    return v;
  };
  for (int t = threadIdx.x; t < wcnt * n2p; t += blockDim.x) {
    const int w = t / n2p, e = t - w * n2p;
    double2 v = make_double2(0.0, 0.0);
    if (e < ntp) v = gp(e, k0 + w);
    int i2 = (mirror_base - e) % n2p;
    if (i2 < 0) i2 += n2p;
    if (i2 < ntp) v = zadd(v, gp(i2, k0 + w + half));
    a[w * lmax + e] = v;
  }
  __syncthreads();
  double2* X = lines64(a, b, wcnt, lmax, pp, 0);
  double2* Y = (X == a) ? b : a;
  resample64_adj(X, Y, wcnt, lmax, lmax, n2c, n2p, ph, 1.0 / n2c);
  __syncthreads();
  double2* Z = lines64(Y, X, wcnt, lmax, pc, 1);
  double2* g = gG + ((int64_t)c * F + f) * ntc * (int64_t)npp;
  for (int t = threadIdx.x; t < ntc * wcnt; t += blockDim.x) {
    const int j = t / wcnt, w = t - j * wcnt;
    g[(int64_t)j * npp + k0 + w] = Z[w * lmax + j];
    g[(int64_t)(ntc - 1 - j) * npp + k0 + w + half] = Z[w * lmax + ntc + j];
  }
}

// Adjoint of phi64 on rows: [plane][ntc][npp] -> [plane][ntc][npc].
__global__ __launch_bounds__(kT) void phi_adj64_kernel(const double2* __restrict__ in,
                                                       double2* __restrict__ out, int ntc,
                                                       int npc, int npp, int rows_per_block,
                                                       wf::Fft64Plan pp0, wf::Fft64Plan pc0) {
  extern __shared__ double2 sm64[];
  const int lmax = npc > npp ? npc : npp;
  double2* a = sm64;
  double2* b = sm64 + rows_per_block * lmax;
  const wf::Fft64Plan pp = plan_in_lds(pp0, b + rows_per_block * lmax);
  const wf::Fft64Plan pc = plan_in_lds(pc0, b + rows_per_block * lmax + pp0.n);
  const int rblocks = (ntc + rows_per_block - 1) / rows_per_block;
  const int64_t plane = blockIdx.x / rblocks;
  const int row0 = (int)(blockIdx.x - plane * rblocks) * rows_per_block;
  const int nrows = min(rows_per_block, ntc - row0);
  const double2* src = in + (plane * ntc + row0) * (int64_t)npp;
  for (int t = threadIdx.x; t < nrows * npp; t += blockDim.x) {
    const int r = t / npp, k = t - r * npp;
    a[r * lmax + k] = src[(int64_t)r * npp + k];
  }
  __syncthreads();
  double2* X = lines64(a, b, nrows, lmax, pp, 0);
  double2* Y = (X == a) ? b : a;
  resample64_adj(X, Y, nrows, lmax, lmax, npc, npp, nullptr, 1.0 / npc);
  __syncthreads();
  double2* Z = lines64(Y, X, nrows, lmax, pc, 1);
  double2* dst = out + (plane * ntc + row0) * (int64_t)npc;
  for (int t = threadIdx.x; t < nrows * npc; t += blockDim.x) {
    const int r = t / npc, k = t - r * npc;
    dst[(int64_t)r * npc + k] = Z[r * lmax + k];
  }
}

// Leaf back-projection: Q[n][f][0..3] = sum_d conj(P[d][n]) {1, s_x, s_y, s_z} g[leaf][f][d].
__global__ __launch_bounds__(kT) void leaf_adj64_kernel(wf::LeafAdj64Args a) {
  __shared__ double3 ss[kTile];
  __shared__ double2 sg[kTile][kFB];
  __shared__ double2 red[kT][4];
  const int nfb = (a.F + kFB - 1) / kFB;
  const int leaf = (int)(blockIdx.x / nfb);
  const int fb = (int)(blockIdx.x % nfb) * kFB;
  const int D = a.nt * a.np;
  const int b = a.leaf_begin[leaf], e = a.leaf_end[leaf];
  double c[3];
  leaf_centre64(a.leaf_key[leaf], a.depth, a.box_lo, a.leaf_size, c);
  const double k_rev = a.k0 / WF_TWO_PI;
  // a leaf of PB < 256 points gets S = min(8, 256 / PB) threads per point,
  // each summing every S-th direction; the S partial sums are added in a
  // fixed order at the end (deterministic)
  const int P = e - b;
  const int PB = P >= kT ? kT : (P > 0 ? P : 1);
  const int S = min(8, kT / PB);
  const int pi = (int)threadIdx.x % PB, sl = (int)threadIdx.x / PB;
  const bool active = sl < S;
  for (int pb = b; pb < e; pb += PB) {
    const int i = pb + pi;
    const bool valid = active && i < e;
    double r[3] = {0.0, 0.0, 0.0};
    if (valid) {
      r[0] = ((double)a.px[i] - c[0]) * k_rev;
      r[1] = ((double)a.py[i] - c[1]) * k_rev;
      r[2] = ((double)a.pz[i] - c[2]) * k_rev;
    }
    double2 acc[kFB][4];
#pragma unroll
    for (int f = 0; f < kFB; ++f)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[f][j] = make_double2(0.0, 0.0);
    for (int d0 = 0; d0 < D; d0 += kTile) {
      const int cnt = min(kTile, D - d0);
      __syncthreads();
      for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
        double sd[3];
        leaf_dir64(d0 + t, a.nt, a.np, sd);
        ss[t] = make_double3(sd[0], sd[1], sd[2]);
      }
      for (int t = threadIdx.x; t < cnt * kFB; t += blockDim.x) {
        const int dd = t / kFB, f = t - dd * kFB;
        sg[dd][f] = (fb + f < a.F) ? a.g[((int64_t)leaf * a.F + fb + f) * D + d0 + dd]
                                   : make_double2(0.0, 0.0);
      }
      __syncthreads();
      if (valid) {
#pragma unroll 2
        for (int dd = sl; dd < cnt; dd += S) {  // two independent phase chains in flight
          const double3 sv = ss[dd];
          const double2 ph = zcis_rev(sv.x * r[0] + sv.y * r[1] + sv.z * r[2]);
#pragma unroll
          for (int f = 0; f < kFB; ++f) {
            const double2 g = sg[dd][f];
            // conj(P) g
            const double2 v = make_double2(ph.x * g.x + ph.y * g.y, ph.x * g.y - ph.y * g.x);
            acc[f][0] = zadd(acc[f][0], v);
            acc[f][1] = zadd(acc[f][1], zscale(v, sv.x));
            acc[f][2] = zadd(acc[f][2], zscale(v, sv.y));
            acc[f][3] = zadd(acc[f][3], zscale(v, sv.z));
          }
        }
      }
    }
    // add the slices' partial sums, slice 0 first
#pragma unroll
    for (int f = 0; f < kFB; ++f) {
      __syncthreads();
#pragma unroll
      for (int j = 0; j < 4; ++j) red[threadIdx.x][j] = acc[f][j];
      __syncthreads();
      if (valid && sl == 0 && fb + f < a.F) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          double2 q = red[pi][j];
          for (int u = 1; u < S; ++u) q = zadd(q, red[pi + u * PB][j]);
          a.Q[((int64_t)i * a.F + fb + f) * 4 + j] = q;
        }
      }
    }
  }
}

// LDS budget of the line kernels: two ping-pong buffers of this many double2
constexpr int kBatch64 = 2048;          // line buffers + tables: 64 KB
constexpr int kSlots64 = 16;            // parent slots per thread in theta64
constexpr size_t kLds64 = 80 * 1024;  // lines + tables (a table of one long line may exceed 64 KB)

// Lines per block: two buffers of kBatch64 samples, less the tables
// (extra samples: twiddles, shift factors) kept beside them.
int rows_for(int lmax, int ntc, int extra) {
  int rows = (kBatch64 - extra / 2) / lmax;
  if (rows < 1) rows = 1;
  if (rows > ntc) rows = ntc;
  return rows;
}

int cols_for(int lmax, int n2p, int half, int extra) {
  int W = std::min((kBatch64 - extra / 2) / lmax, (kT * kSlots64) / n2p);
  if (W < 1) W = 1;
  if (W > half) W = half > 0 ? half : 1;
  return W;
}

}  // namespace

namespace wf {

hipError_t launch_leaf64(const Leaf64Args& a, hipStream_t st) {
  if (a.nleaves <= 0) return hipSuccess;
  if (a.F % kFB != 0) return hipErrorInvalidValue;
  const int D = a.nt * a.np;
  const int64_t tiles = (D + kT - 1) / kT;
  const int64_t nblk = tiles * (a.F / kFB) * a.nleaves;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(leaf64_kernel, dim3((unsigned)nblk), dim3(kT), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_phi64(const double2* in, double2* out, int64_t planes, int ntc, int npc,
                        int npp, const Fft64Plan& pc, const Fft64Plan& pp, hipStream_t st) {
  if (planes <= 0) return hipSuccess;
  const int lmax = npc > npp ? npc : npp;
  if (lmax > 2 * kBatch64) return hipErrorInvalidValue;
  const int extra = pc.n + pp.n;  // twiddle tables
  const int rows = rows_for(lmax, ntc, extra);
  const size_t sh = ((size_t)2 * rows * lmax + extra) * sizeof(double2);
  if (sh > kLds64) return hipErrorInvalidValue;
  {
    const hipError_t e = hipFuncSetAttribute((const void*)phi64_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  const int64_t nblk = (int64_t)((ntc + rows - 1) / rows) * planes;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(phi64_kernel, dim3((unsigned)nblk), dim3(kT), sh, st, in, out, ntc, npc, npp,
                     rows, pc, pp);
  return hipGetLastError();
}

hipError_t launch_theta64(const double2* G, int F, int ntc, int npp, int ntp, int nparents,
                          const int32_t* cb, const int32_t* ce, const uint32_t* child_key,
                          const double2* T, double shift, int mirror_base, const Fft64Plan& pc,
                          const Fft64Plan& pp, double2* out, hipStream_t st) {
  if (nparents <= 0) return hipSuccess;
  const int n2c = 2 * ntc, n2p = 2 * ntp;
  const int lmax = n2c > n2p ? n2c : n2p;
  const int half = npp / 2;
  const int extra = n2c + 1 + pc.n + pp.n;  // shift factors, twiddle tables
  const int W = cols_for(lmax, n2p, half, extra);
  if (W * n2p > kT * kSlots64) return hipErrorInvalidValue;
  const size_t sh = ((size_t)2 * W * lmax + extra) * sizeof(double2);
  if (sh > kLds64) return hipErrorInvalidValue;
  const int64_t nblk = (int64_t)((half + W - 1) / W) * F * nparents;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  {
    const hipError_t e = hipFuncSetAttribute((const void*)theta64_kernel<kSlots64>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((theta64_kernel<kSlots64>), dim3((unsigned)nblk), dim3(kT), sh, st, G, F, ntc,
                     npp, ntp, W, cb, ce, child_key, T, shift, mirror_base, pc, pp, out);
  return hipGetLastError();
}

hipError_t launch_trans64(double2* T, int ntp, int npp, double theta0, const double h[3], double k0,
                          hipStream_t st) {
  const int64_t tot = 8LL * ntp * npp;
  int64_t blocks = (tot + kT - 1) / kT;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) return hipSuccess;
  hipLaunchKernelGGL(trans64_kernel, dim3((unsigned)blocks), dim3(kT), 0, st, T, ntp, npp, theta0,
                     h[0], h[1], h[2], k0 / WF_TWO_PI);
  return hipGetLastError();
}

hipError_t launch_theta_adj64(const double2* gP, int F, int ntc, int npp, int ntp, int nchild,
                              const int32_t* parent_idx, const uint32_t* child_key,
                              const double2* T, double shift, int mirror_base,
                              const Fft64Plan& pp, const Fft64Plan& pc, double2* gG,
                              hipStream_t st) {
  if (nchild <= 0) return hipSuccess;
  const int n2c = 2 * ntc, n2p = 2 * ntp;
  const int lmax = n2c > n2p ? n2c : n2p;
  const int half = npp / 2;
  const int extra = n2c + 1 + pc.n + pp.n;  // shift factors, twiddle tables
  const int W = cols_for(lmax, n2p, half, extra);
  const size_t sh = ((size_t)2 * W * lmax + extra) * sizeof(double2);
  if (sh > kLds64) return hipErrorInvalidValue;
  const int64_t nblk = (int64_t)((half + W - 1) / W) * F * nchild;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  {
    const hipError_t e = hipFuncSetAttribute((const void*)theta_adj64_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(theta_adj64_kernel, dim3((unsigned)nblk), dim3(kT), sh, st, gP, F, ntc, npp,
                     ntp, W, parent_idx, child_key, T, shift, mirror_base, pp, pc, gG);
  return hipGetLastError();
}

hipError_t launch_phi_adj64(const double2* in, double2* out, int64_t planes, int ntc, int npc,
                            int npp, const Fft64Plan& pp, const Fft64Plan& pc, hipStream_t st) {
  if (planes <= 0) return hipSuccess;
  const int lmax = npc > npp ? npc : npp;
  const int extra = pc.n + pp.n;  // twiddle tables
  const int rows = rows_for(lmax, ntc, extra);
  const size_t sh = ((size_t)2 * rows * lmax + extra) * sizeof(double2);
  if (sh > kLds64) return hipErrorInvalidValue;
  {
    const hipError_t e = hipFuncSetAttribute((const void*)phi_adj64_kernel,
                                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  const int64_t nblk = (int64_t)((ntc + rows - 1) / rows) * planes;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(phi_adj64_kernel, dim3((unsigned)nblk), dim3(kT), sh, st, in, out, ntc, npc,
                     npp, rows, pp, pc);
  return hipGetLastError();
}

hipError_t launch_leaf_adj64(const LeafAdj64Args& a, hipStream_t st) {
  if (a.nleaves <= 0) return hipSuccess;
  const int nfb = (a.F + kFB - 1) / kFB;
  const int64_t nblk = (int64_t)a.nleaves * nfb;
  if (nblk > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(leaf_adj64_kernel, dim3((unsigned)nblk), dim3(kT), 0, st, a);
  return hipGetLastError();
}

}  // namespace wf
