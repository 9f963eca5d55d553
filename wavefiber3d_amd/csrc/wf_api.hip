// wf_api.hip -- C-ABI of the MI355X PO solver: handle, geometry/octree cache,
// per-wavelength solve orchestration (drivers treePO.cu:21-992,
// sphere.cu:21-966 restructured as a library).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wf_po.h"
#include "wf_adjoint.h"
#include "wf_common.h"
#include "wf_fft.h"
#include "wf_interp.h"
#include "wf_octree.h"
#include "wf_tree64.h"

namespace {

thread_local char g_err[1024] = "";

wf_status set_err(wf_status s, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return s;
}

#define WF_HIP(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return set_err(WF_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));     \
  } while (0)

#define WF_TRY(expr)                  \
  do {                                \
    wf_status s_ = (expr);            \
    if (s_ != WF_OK) return s_;       \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
};

struct LevelDev {
  int count = 0;
  uint32_t* key = nullptr;   // level-local node ids (level order)
  int32_t* begin = nullptr;  // first sorted point
  int32_t* end = nullptr;
  int32_t* cb = nullptr;     // children range in the next level's list
  int32_t* ce = nullptr;
  int32_t* parent = nullptr; // index of the parent in the previous level's list
  int32_t* span = nullptr;   // {0, count}: the whole level as one child range
};

struct FftEntry {
  FftPlanDev plan;
  float2* tw = nullptr;
};

struct Fft2Entry {
  Fft2Plan plan;
  float2* tw = nullptr;
};

// Split-BF16 interpolation matrix of one (input length, output length, shift).
struct MatEntry {
  uint16_t* hi = nullptr;
  uint16_t* lo = nullptr;
  int rows_pad = 0, kpad = 0;
};

// A captured solve; valid while the workspace it was recorded on has not
// moved (ws_gen) and the call signature matches.
struct GraphCache {
  bool valid = false;
  uint64_t sig = 0;
  uint64_t ws_gen = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    valid = false;
  }
};

}  // namespace

struct wf_handle {
  wf_opts opts;
  int device = 0;
  size_t bytes_in_use = 0;
  std::map<std::string, DevBuf> bufs;
  std::map<int, FftEntry> ffts;
  std::map<std::pair<std::pair<int, int>, uint64_t>, MatEntry> mats;
  std::map<int, Fft2Entry> ffts2;
  std::map<int, wf::Fft64Plan> ffts64;  // FP64 twiddles (owned: plan.tw)
  std::map<std::pair<int, uint64_t>, float2*> facs;
  // geometry
  bool has_geom = false;
  int64_t n = 0;
  int depth = 0;           // depth used by the current solve (<= depth_max)
  int depth_max = 0;       // depth the octree was built to
  bool auto_depth = false; // WF_DEPTH_AUTO: per-group depth from the cost model
  int force_depth = -1;    // autotuning: depth forced for every group
  wf_tune tune;
  // depth measured fastest by wf_autotune for this geometry (geom_gen) at the
  // sampling wavelength tuned_lam; other wavelengths use the cost model
  uint64_t tuned_gen = 0;
  double tuned_lam = 0;
  int tuned_depth = -1;
  float boxf[6] = {0, 0, 0, 0, 0, 0};
  double box[6] = {0, 0, 0, 0, 0, 0};
  float* pts = nullptr;     // sorted SoA: x y z nx ny nz area (7 n)
  int32_t* perm = nullptr;  // sorted position -> original index
  uint32_t* keys = nullptr; // sorted keys
  std::vector<LevelDev> levels;
  // incidence upload (pinned staging)
  WfIncDev* inc_host = nullptr;
  int inc_host_cap = 0;
  hipEvent_t inc_evt = nullptr;
  // per-launch timing: events from a reusable pool, summed lazily
  std::vector<hipEvent_t> evpool;
  int ev_used = 0;
  struct Rec {
    int kind, e0, e1;
    double work;
  };
  std::vector<Rec> recs;
  int ev_t0 = -1, ev_t1 = -1;
  bool stats_final = true;
  wf_stats stats;
  // hipGraph replay of wf_solve (opts.use_graph)
  bool capturing = false;
  uint64_t geom_gen = 0;
  uint64_t ws_gen = 1;  // bumped whenever a workspace buffer moves
  GraphCache gcache;
  // concurrent chunks of wf_solve (opts.solve_streams): worker handles bound
  // to this handle's geometry, each with its own workspace and stream
  std::vector<wf_handle*> subs;
  std::vector<hipStream_t> sub_st;
  std::vector<hipEvent_t> sub_ev;  // [0]: fork on the caller's stream, [1 + k]: join of worker k
  bool sequential = false;         // wf_autotune: clean per-kernel times
};

namespace {

wf_status dalloc(wf_handle* h, const char* name, size_t bytes, void** out) {
  DevBuf& b = h->bufs[name];
  if (bytes == 0) bytes = 16;
  if (b.cap >= bytes) {
    *out = b.p;
    return WF_OK;
  }
  h->ws_gen++;  // captured graphs may hold the old address
  if (b.p) {
    hipFree(b.p);
    h->bytes_in_use -= b.cap;
    b.p = nullptr;
    b.cap = 0;
  }
  const size_t want = bytes + bytes / 8;  // headroom against regrowth
  if (h->opts.workspace_cap_bytes > 0 && h->bytes_in_use + want > (size_t)h->opts.workspace_cap_bytes)
    return set_err(WF_ERR_OOM, "workspace cap %lld bytes exceeded (buffer %s needs %zu)",
                   (long long)h->opts.workspace_cap_bytes, name, want);
  hipError_t e = hipMalloc(&b.p, want);
  if (e != hipSuccess) {
    b.p = nullptr;
    hipGetLastError();
    return set_err(WF_ERR_OOM, "hipMalloc(%zu) for %s failed: %s", want, name, hipGetErrorString(e));
  }
  b.cap = want;
  h->bytes_in_use += want;
  *out = b.p;
  return WF_OK;
}

template <typename T>
wf_status dalloc_t(wf_handle* h, const char* name, size_t count, T** out) {
  void* p = nullptr;
  wf_status s = dalloc(h, name, count * sizeof(T), &p);
  *out = static_cast<T*>(p);
  return s;
}

bool cancelled(wf_handle* h) { return h->opts.cancel_flag && *h->opts.cancel_flag != 0; }

// 13-smooth lengths: radices 4, 2, 3, 5, 7, 11, 13 of the in-house FFT.
bool factor_len(int n, FftPlanDev& p) {
  if (n < 1) return false;
  p.n = n;
  p.nstages = 0;
  p.radix_packed = 0;
  int m = n;
  const int rad[] = {4, 2, 3, 5, 7, 11, 13};
  for (int r : rad) {
    while (m % r == 0) {
      if (p.nstages >= WF_FFT_MAX_STAGES) return false;
      p.radix_packed |= (uint64_t)r << (4 * p.nstages++);
      m /= r;
    }
  }
  return m == 1;
}

bool is_smooth(int n) {
  FftPlanDev p;
  return factor_len(n, p);
}

// Internal level grids use 5-smooth lengths (radix 2, 3, 4, 5 butterflies);
// 7, 11 and 13 are only needed for caller-chosen output grids.
bool is_smooth5(int n) {
  for (int r : {2, 3, 5})
    while (n > 1 && n % r == 0) n /= r;
  return n == 1;
}
int smooth_up(int n) {
  while (!is_smooth5(n)) ++n;
  return n;
}
int smooth_even_up(int n) {
  n = smooth_up(n);
  while (n % 2) n = smooth_up(n + 1);
  return n;
}

wf_status get_fft(wf_handle* h, int n, FftPlanDev* out) {
  auto it = h->ffts.find(n);
  if (it != h->ffts.end()) {
    *out = it->second.plan;
    return WF_OK;
  }
  FftEntry e;
  if (!factor_len(n, e.plan))
    return set_err(WF_ERR_UNSUPPORTED, "FFT length %d has a prime factor above 13", n);
  std::vector<float2> tw(n);
  for (int m = 0; m < n; ++m) {
    const double a = -2.0 * WF_PI * (double)m / (double)n;
    tw[m] = make_float2((float)std::cos(a), (float)std::sin(a));
  }
  WF_HIP(hipMalloc(&e.tw, sizeof(float2) * n));
  WF_HIP(hipMemcpy(e.tw, tw.data(), sizeof(float2) * n, hipMemcpyHostToDevice));
  e.plan.tw = e.tw;
  h->ffts[n] = e;
  *out = e.plan;
  return WF_OK;
}

// Interpolation matrix from nc samples to np samples with output origin
// shifted by `shift` (radians of the period-2pi parameter), cached per handle.
// Interpolation matrix nc -> np samples (gc: the theta matrix of the MFMA
// pass, nc = 2 ntc, columns in the phi pass's great-circle row order).
// FP64 Stockham plan of length n (radices as get_fft, twiddles in FP64).
wf_status get_fft64(wf_handle* h, int n, wf::Fft64Plan* out) {
  auto it = h->ffts64.find(n);
  if (it != h->ffts64.end()) {
    *out = it->second;
    return WF_OK;
  }
  FftPlanDev f;
  if (!factor_len(n, f))
    return set_err(WF_ERR_UNSUPPORTED, "FFT length %d has a prime factor above 13", n);
  std::vector<double2> tw(n);
  for (int m = 0; m < n; ++m) {
    const double a = -2.0 * WF_PI * (double)m / (double)n;
    tw[m] = make_double2(std::cos(a), std::sin(a));
  }
  double2* d = nullptr;
  WF_HIP(hipMalloc(&d, sizeof(double2) * n));
  WF_HIP(hipMemcpy(d, tw.data(), sizeof(double2) * n, hipMemcpyHostToDevice));
  wf::Fft64Plan p;
  p.n = n;
  p.nstages = f.nstages;
  p.radix_packed = f.radix_packed;
  p.tw = d;
  h->ffts64[n] = p;
  *out = p;
  return WF_OK;
}

wf_status get_mat(wf_handle* h, int nc, int np, double shift, MatEntry* out, bool gc = false,
                  bool flip = false) {
  uint64_t sb;
  memcpy(&sb, &shift, sizeof sb);
  // keys: nc for phi matrices, -nc for theta matrices, -nc - 2^20 for the
  // sign-flipped theta matrices of transverse components
  const auto key = std::make_pair(std::make_pair(gc ? (flip ? -nc - (1 << 20) : -nc) : nc, np), sb);
  auto it = h->mats.find(key);
  if (it != h->mats.end()) {
    *out = it->second;
    return WF_OK;
  }
  std::vector<uint16_t> hi, lo;
  MatEntry e;
  if (gc)
    wf::build_theta_matrix_gc(nc / 2, np, shift, hi, lo, e.rows_pad, e.kpad, flip);
  else
    wf::build_interp_matrix(nc, np, shift, hi, lo, e.rows_pad, e.kpad);
  WF_HIP(hipMalloc(&e.hi, sizeof(uint16_t) * hi.size()));
  WF_HIP(hipMalloc(&e.lo, sizeof(uint16_t) * lo.size()));
  WF_HIP(hipMemcpy(e.hi, hi.data(), sizeof(uint16_t) * hi.size(), hipMemcpyHostToDevice));
  WF_HIP(hipMemcpy(e.lo, lo.data(), sizeof(uint16_t) * lo.size(), hipMemcpyHostToDevice));
  h->mats[key] = e;
  *out = e;
  return WF_OK;
}

// FFT plan of the large-level kernels: radices 4, then 3, 5, 2; per-stage
// twiddles exp(-2 pi i r k / (Ns R)) in FP64, stored float.
wf_status get_fft2(wf_handle* h, int n, Fft2Plan* out) {
  auto it = h->ffts2.find(n);
  if (it != h->ffts2.end()) {
    *out = it->second.plan;
    return WF_OK;
  }
  Fft2Entry e;
  e.plan.n = n;
  e.plan.nstages = 0;
  e.plan.radix_packed = 0;
  memset(e.plan.mag_ns, 0, sizeof e.plan.mag_ns);
  int m = n;
  std::vector<int> rad;
  // odd radices first, then the power of two 2^a: 4s for even a, for odd a
  // one 8 (a >= 3 and n <= 512) or one 2 followed by 4s.  Even lengths end on
  // radix 8, 4 or 2
  // (the wave-private kernels are compiled per last radix); n = 4 runs as two
  // radix-2 stages (those kernels need at least two stages).
  for (int r : {5, 3})
    while (m % r == 0) {
      rad.push_back(r);
      m /= r;
    }
  int a2 = 0;
  while (m % 2 == 0) {
    m /= 2;
    ++a2;
  }
  if (n == 4 || n == 8) {
    rad.push_back(n / 2 == 2 ? 2 : 4);
    rad.push_back(2);
  } else {
    if (a2 % 2 == 1 && a2 >= 3 && n <= 512) {  // the wave kernels' radix-8 bound
      rad.push_back(8);
      a2 -= 3;
    } else if (a2 % 2 == 1) {
      rad.push_back(2);
      a2 -= 1;
    }
    for (; a2 >= 2; a2 -= 2) rad.push_back(4);
  }
  if (m != 1 || n < 1 || (int)rad.size() > WF_FFT2_MAX_STAGES)
    return set_err(WF_ERR_UNSUPPORTED, "FFT length %d is not 5-smooth", n);
  std::vector<float2> tw;
  int Ns = 1;
  for (size_t s = 0; s < rad.size(); ++s) {
    const int R = rad[s];
    e.plan.radix_packed |= (uint64_t)R << (4 * s);
    e.plan.mag_ns[s] = fmagic((uint32_t)Ns);
    for (int k = 0; k < Ns; ++k)
      for (int r = 1; r < R; ++r) {
        const double a = -2.0 * WF_PI * (double)r * k / ((double)Ns * R);
        tw.push_back(make_float2((float)std::cos(a), (float)std::sin(a)));
      }
    Ns *= R;
  }
  e.plan.nstages = (int)rad.size();
  e.plan.twn = (int)tw.size();
  e.plan.mag_n = fmagic((uint32_t)n);
  WF_HIP(hipMalloc(&e.tw, sizeof(float2) * std::max<size_t>(tw.size(), 1)));
  if (!tw.empty())
    WF_HIP(hipMemcpy(e.tw, tw.data(), sizeof(float2) * tw.size(), hipMemcpyHostToDevice));
  e.plan.tw = e.tw;
  h->ffts2[n] = e;
  *out = e.plan;
  return WF_OK;
}

// Resample factors w_f exp(i f shift) / nc for f = -(nc/2) .. band top
// (nc + 1 entries; the Nyquist pair of an even nc gets 1/2 each).
wf_status get_fac(wf_handle* h, int nc, double shift, const float2** out) {
  uint64_t sb;
  memcpy(&sb, &shift, sizeof sb);
  const auto key = std::make_pair(nc, sb);
  auto it = h->facs.find(key);
  if (it != h->facs.end()) {
    *out = it->second;
    return WF_OK;
  }
  std::vector<float2> fac(nc + 1, make_float2(0.f, 0.f));
  const int fmin = -(nc / 2), fmax = (nc - 1) / 2 + ((nc % 2 == 0) ? 1 : 0);
  for (int f = fmin; f <= fmax; ++f) {
    double w = 1.0 / nc;
    if (nc % 2 == 0 && (f == nc / 2 || f == -(nc / 2))) w *= 0.5;
    fac[f - fmin] = make_float2((float)(w * std::cos(f * shift)), (float)(w * std::sin(f * shift)));
  }
  float2* d = nullptr;
  WF_HIP(hipMalloc(&d, sizeof(float2) * fac.size()));
  WF_HIP(hipMemcpy(d, fac.data(), sizeof(float2) * fac.size(), hipMemcpyHostToDevice));
  h->facs[key] = d;
  *out = d;
  return WF_OK;
}

// Incidence parameters in FP64 (treePO.cu:501-509, :580-601).
WfIncDev make_inc(const wf_incidence& in) {
  WfIncDev d;
  memset(&d, 0, sizeof d);
  const double ti = in.theta, pi_ = in.phi;
  d.w[0] = std::cos(ti) * std::cos(pi_);
  d.w[1] = std::cos(ti) * std::sin(pi_);
  d.w[2] = std::sin(ti);
  const double a[3] = {-std::sin(ti) * std::cos(pi_), -std::sin(ti) * std::sin(pi_), std::cos(ti)};
  const double b[3] = {std::sin(pi_), -std::cos(pi_), 0.0};
  const double c[3] = {-std::sin(pi_), std::cos(pi_), 0.0};
  for (int q = 0; q < 3; ++q) {
    if (in.pol == 'M') {
      d.E[q] = a[q];
      d.H[q] = b[q];
    } else {
      d.E[q] = c[q];
      d.H[q] = a[q];
    }
  }
  const double mu0 = 4e-7 * WF_PI, eps0 = 8.8541878e-12, c0 = 299792458.0;
  d.k0 = 2.0 * WF_PI / in.lambda_m;
  d.z0 = std::sqrt(mu0 / eps0);
  d.sigma = (std::isfinite(in.beam_sigma_m) && in.beam_sigma_m > 0) ? in.beam_sigma_m : 0.0;
  d.eta_re = in.eta_re;
  d.eta_im = in.eta_im;
  const double omega = d.k0 * c0;
  // e^{-i k0 D}/D, D = 1 m; E_f = -i w mu0/(4 pi) g, H_f = -i w eps0/(4 pi) g
  const double gr = std::cos(-d.k0), gi = std::sin(-d.k0);
  const double fe = omega * mu0 / (4 * WF_PI), fh = omega * eps0 / (4 * WF_PI);
  d.ef_re = fe * gi;   // -i * fe * (gr + i gi) = fe*gi - i fe*gr
  d.ef_im = -fe * gr;
  d.hf_re = fh * gi;
  d.hf_im = -fh * gr;
  d.mode = in.pol;
  return d;
}

wf_status check_inc(const wf_incidence* inc, int n_inc) {
  if (!inc || n_inc <= 0) return set_err(WF_ERR_INVALID_ARG, "incidences: null or n_inc <= 0");
  for (int i = 0; i < n_inc; ++i) {
    const wf_incidence& t = inc[i];
    if (!(t.pol == 'M' || t.pol == 'E'))
      return set_err(WF_ERR_INVALID_ARG, "incidence %d: pol must be 'M' or 'E'", i);
    if (!(std::isfinite(t.lambda_m) && t.lambda_m > 0))
      return set_err(WF_ERR_INVALID_ARG, "incidence %d: lambda_m must be finite and > 0", i);
    if (!(std::isfinite(t.theta) && std::isfinite(t.phi)))
      return set_err(WF_ERR_INVALID_ARG, "incidence %d: theta/phi must be finite", i);
    if (!(std::isfinite(t.eta_re) && std::isfinite(t.eta_im)) || t.eta_im < 0 ||
        (t.eta_re == -1.0 && t.eta_im == 0.0) || (t.eta_re == 0.0 && t.eta_im == 0.0))
      return set_err(WF_ERR_INVALID_ARG, "incidence %d: eta must be finite with kappa >= 0", i);
  }
  return WF_OK;
}

wf_status upload_incs(wf_handle* h, const wf_incidence* inc, int n_inc, WfIncDev** dev,
                      hipStream_t st, const char* buf = "inc") {
  if (h->inc_host_cap < n_inc) {
    if (h->inc_host) {
      hipEventSynchronize(h->inc_evt);
      hipHostFree(h->inc_host);
    }
    WF_HIP(hipHostMalloc((void**)&h->inc_host, sizeof(WfIncDev) * n_inc));
    h->inc_host_cap = n_inc;
  } else {
    WF_HIP(hipEventSynchronize(h->inc_evt));
  }
  for (int i = 0; i < n_inc; ++i) h->inc_host[i] = make_inc(inc[i]);
  WF_TRY(dalloc_t(h, buf, n_inc, dev));
  WF_HIP(hipMemcpyAsync(*dev, h->inc_host, sizeof(WfIncDev) * n_inc, hipMemcpyHostToDevice, st));
  WF_HIP(hipEventRecord(h->inc_evt, st));
  return WF_OK;
}

struct Grid {
  int nt, np;
};

// Level sampling.  The reference rule (treePO.cu:342-375) is
// N = ceil(k d + beta (k d)^(1/3)) with d = 0.5^i * (extent * rho).  Here the
// extent is the node's own bound: theta samples the great circle through the
// poles (band limit k R, R = node half-diagonal, 2 pi extension -> Nt >= L),
// phi samples the azimuth (band limit k R_xy, Np >= 2 L), rounded up to FFT
// lengths (Np even for the extension).
Grid level_grid(const wf_handle* h, int level, double lambda_s) {
  const double ks = 2.0 * WF_PI / lambda_s;
  double hx = 0.5 * (h->box[3] - h->box[0]) / std::ldexp(1.0, level);
  double hy = 0.5 * (h->box[4] - h->box[1]) / std::ldexp(1.0, level);
  double hz = 0.5 * (h->box[5] - h->box[2]) / std::ldexp(1.0, level);
  const double R = std::sqrt(hx * hx + hy * hy + hz * hz);
  const double Rxy = std::sqrt(hx * hx + hy * hy);
  const double beta = h->opts.beta;
  Grid g;
  if (h->opts.ref_dtheta_m > 0 && h->opts.ref_dphi_m > 0) {
    // the reference's rule (treePO.cu:346-351): d in float, the count in double
    auto N = [&](double d0) {
      const float d = (float)(std::ldexp(1.0, -level) * d0);
      const double kd = ks * d;
      return (int)std::ceil(kd + beta * std::pow(kd, 1.0 / 3.0));
    };
    g.nt = smooth_up(std::max(4, N(h->opts.ref_dtheta_m)));
    g.np = smooth_even_up(std::max(8, N(h->opts.ref_dphi_m)));
    return g;
  }
  const double bt = h->tune.beta > 0 ? h->tune.beta : beta;  // the tuned constant
  auto L = [&](double r) { return ks * r + bt * std::cbrt(ks * r); };
  g.nt = smooth_up(std::max(4, (int)std::ceil(L(R))));
  g.np = smooth_even_up(std::max(8, (int)std::ceil(2.0 * L(Rxy))));
  return g;
}

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Cost-model defaults (measured on MI355X with the BF16x3 quad leaf kernel
// and the fused aggregation; wf_autotune refits them per geometry).
void default_tune(wf_tune* t) {
  t->leaf_s_per_pair = 6.0e-14;
  t->agg_s_per_byte = 4.0e-13;
  t->level_s = 2.0e-6;
  t->leaf_quad = 1;
  t->chunk_solves = 4;
  t->agg_kernel = 0;
  t->agg_fft_rows = 600;
  t->calibrated = 0;
  t->agg_fused = 2;
  t->theta_tile = 0;
  t->phi_rows = 0;
  t->beta = 0.0;
}

// Deepest useful level of a WF_DEPTH_AUTO build: about one octant digit per
// factor 8 of points (ceil(log8 n)), at least 1, at most 10 (30-bit keys).
int auto_depth_max(int64_t n) {
  int d = 0;
  double m = 1.0;
  while (m < (double)n && d < 10) {
    m *= 8.0;
    ++d;
  }
  return std::max(1, d);
}

wf_status set_device(wf_handle* h) {
  if (h->device >= 0) WF_HIP(hipSetDevice(h->device));
  return WF_OK;
}

// ---------------------------------------------------------------- timing --
wf_status prof_event(wf_handle* h, hipStream_t st, int* idx) {
  if (h->capturing) {
    *idx = -1;
    return WF_OK;
  }
  if (h->ev_used == (int)h->evpool.size()) {
    hipEvent_t e;
    WF_HIP(hipEventCreate(&e));
    h->evpool.push_back(e);
  }
  *idx = h->ev_used++;
  WF_HIP(hipEventRecord(h->evpool[*idx], st));
  return WF_OK;
}

void prof_reset(wf_handle* h) {
  h->ev_used = 0;
  h->recs.clear();
  h->ev_t0 = h->ev_t1 = -1;
  memset(&h->stats, 0, sizeof h->stats);
  h->stats_final = false;
  for (wf_handle* w : h->subs) prof_reset(w);
}

// Runs `launch` (returning hipError_t) between two events and books it.
template <typename Fn>
wf_status timed(wf_handle* h, hipStream_t st, int kind, double work, Fn&& launch) {
  if (h->capturing) {  // graph capture: no timing events inside the graph
    hipError_t e = launch();
    if (e != hipSuccess)
      return set_err(WF_ERR_HIP, "kernel launch (kind %d) failed: %s", kind, hipGetErrorString(e));
    return WF_OK;
  }
  int e0, e1;
  WF_TRY(prof_event(h, st, &e0));
  hipError_t e = launch();
  if (e != hipSuccess) return set_err(WF_ERR_HIP, "kernel launch (kind %d) failed: %s", kind, hipGetErrorString(e));
  WF_TRY(prof_event(h, st, &e1));
  h->recs.push_back({kind, e0, e1, work});
  return WF_OK;
}

wf_status prof_finalize(wf_handle* h) {
  if (h->stats_final) return WF_OK;
  if (h->ev_used > 0) WF_HIP(hipEventSynchronize(h->evpool[h->ev_used - 1]));
  for (const auto& r : h->recs) {
    float ms = 0;
    WF_HIP(hipEventElapsedTime(&ms, h->evpool[r.e0], h->evpool[r.e1]));
    h->stats.kern_ms[r.kind] += ms;
    h->stats.kern_launches[r.kind] += 1;
    h->stats.kern_work[r.kind] += r.work;
  }
  if (h->ev_t0 >= 0 && h->ev_t1 >= 0) {
    float ms = 0;
    WF_HIP(hipEventElapsedTime(&ms, h->evpool[h->ev_t0], h->evpool[h->ev_t1]));
    h->stats.ms_total = ms;
  }
  // launches the worker handles made for this handle's last solve
  for (wf_handle* w : h->subs) {
    WF_TRY(prof_finalize(w));
    for (int k = 0; k < WF_K_COUNT; ++k) {
      h->stats.kern_ms[k] += w->stats.kern_ms[k];
      h->stats.kern_launches[k] += w->stats.kern_launches[k];
      h->stats.kern_work[k] += w->stats.kern_work[k];
    }
    h->stats.leaf_pairs += w->stats.leaf_pairs;
    h->stats.agg_transitions += w->stats.agg_transitions;
    h->stats.agg_bytes += w->stats.agg_bytes;
  }
  h->stats_final = true;
  return WF_OK;
}

}  // namespace

// ============================================================== C-ABI =====
extern "C" {

int32_t wf_abi_version(void) { return WF_ABI_VERSION; }

const char* wf_last_error(void) { return g_err; }

void wf_default_opts(wf_opts* o) {
  if (!o) return;
  memset(o, 0, sizeof *o);
  o->struct_size = (int32_t)sizeof(wf_opts);
  o->device = -1;
  o->precision = WF_PREC_BF16X3;
  o->strict_grid = 0;
  o->beta = 4.0;
  o->box_enlarge = 1.1;
  o->lambda_sample_m = 0.0;
  o->has_box = 0;
  o->workspace_cap_bytes = 0;
  o->cancel_flag = nullptr;
  o->field_chunk = 0;
  o->use_graph = 0;
}

wf_status wf_create(wf_handle** out, const wf_opts* opts) {
  if (!out) return set_err(WF_ERR_INVALID_ARG, "wf_create: out is null");
  *out = nullptr;
  wf_opts o;
  wf_default_opts(&o);
  if (opts) {
    if (opts->struct_size != (int32_t)sizeof(wf_opts))
      return set_err(WF_ERR_INVALID_ARG, "wf_create: opts.struct_size %d != %d (ABI mismatch)",
                     opts->struct_size, (int)sizeof(wf_opts));
    o = *opts;
  }
  if (!(o.beta >= 0 && o.beta < 100)) return set_err(WF_ERR_INVALID_ARG, "opts.beta out of range");
  if (o.solve_streams < 0 || o.solve_streams > 16)
    return set_err(WF_ERR_INVALID_ARG, "opts.solve_streams must be in [0, 16]");
  if (!(o.box_enlarge >= 1.0 && o.box_enlarge < 100))
    return set_err(WF_ERR_INVALID_ARG, "opts.box_enlarge must be >= 1");
  if (o.precision != WF_PREC_FP32 && o.precision != WF_PREC_BF16X3 && o.precision != WF_PREC_FP32_VALU)
    return set_err(WF_ERR_INVALID_ARG, "opts.precision unknown");
  if (o.has_box) {
    for (int q = 0; q < 3; ++q)
      if (!(std::isfinite(o.box[q]) && std::isfinite(o.box[q + 3]) && o.box[q + 3] >= o.box[q]))
        return set_err(WF_ERR_INVALID_ARG, "opts.box must be finite with max >= min");
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    hipGetLastError();
    return set_err(WF_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  }
  wf_handle* h = new wf_handle();
  h->opts = o;
  if (o.device >= 0) {
    if (o.device >= ndev) {
      delete h;
      return set_err(WF_ERR_INVALID_ARG, "opts.device %d >= device count %d", o.device, ndev);
    }
    h->device = o.device;
  } else {
    hipGetDevice(&h->device);
  }
  hipSetDevice(h->device);
  if (hipEventCreateWithFlags(&h->inc_evt, hipEventDisableTiming) != hipSuccess) {
    delete h;
    return set_err(WF_ERR_HIP, "hipEventCreate failed");
  }
  memset(&h->stats, 0, sizeof h->stats);
  default_tune(&h->tune);
  *out = h;
  return WF_OK;
}

void wf_destroy(wf_handle* h) {
  if (!h) return;
  hipSetDevice(h->device);
  hipDeviceSynchronize();
  for (wf_handle* w : h->subs) {
    w->has_geom = false;  // the geometry buffers belong to h
    wf_destroy(w);
  }
  for (auto s : h->sub_st) hipStreamDestroy(s);
  for (auto e : h->sub_ev) hipEventDestroy(e);
  for (auto& kv : h->bufs)
    if (kv.second.p) hipFree(kv.second.p);
  for (auto& kv : h->ffts)
    if (kv.second.tw) hipFree(kv.second.tw);
  for (auto& kv : h->mats) {
    if (kv.second.hi) hipFree(kv.second.hi);
    if (kv.second.lo) hipFree(kv.second.lo);
  }
  for (auto& kv : h->ffts2)
    if (kv.second.tw) hipFree(kv.second.tw);
  for (auto& kv : h->ffts64)
    if (kv.second.tw) hipFree(const_cast<double2*>(kv.second.tw));
  for (auto& kv : h->facs)
    if (kv.second) hipFree(kv.second);
  if (h->inc_host) hipHostFree(h->inc_host);
  if (h->inc_evt) hipEventDestroy(h->inc_evt);
  for (auto ev : h->evpool) hipEventDestroy(ev);
  h->gcache.reset();
  delete h;
}

wf_status wf_set_geometry(wf_handle* h, const wf_surface* s, int32_t depth, void* stream) {
  if (!h) return set_err(WF_ERR_INVALID_ARG, "handle is null");
  if (!s) return set_err(WF_ERR_INVALID_ARG, "surface is null");
  if (s->n <= 0) return set_err(WF_ERR_INVALID_ARG, "surface.n must be > 0 (got %lld)", (long long)s->n);
  if (s->n > 0x7fffffffLL) return set_err(WF_ERR_INVALID_ARG, "surface.n too large");
  if ((depth < 0 || depth > 10) && depth != WF_DEPTH_AUTO)
    return set_err(WF_ERR_INVALID_ARG, "depth must be in [0, 10] or WF_DEPTH_AUTO");
  if (!s->x || !s->y || !s->z || !s->nx || !s->ny || !s->nz || !s->area)
    return set_err(WF_ERR_INVALID_ARG, "surface: null array pointer");
  if (s->memspace != WF_MEM_HOST && s->memspace != WF_MEM_DEVICE)
    return set_err(WF_ERR_INVALID_ARG, "surface.memspace unknown");
  WF_TRY(set_device(h));
  hipStream_t st = as_stream(stream);
  h->has_geom = false;
  const int64_t n = s->n;
  const bool auto_depth = depth == WF_DEPTH_AUTO;
  if (auto_depth) depth = auto_depth_max(n);
  // raw device copies (host input) or the caller's device arrays
  const float* src[7] = {s->x, s->y, s->z, s->nx, s->ny, s->nz, s->area};
  if (s->memspace == WF_MEM_HOST) {
    float* raw = nullptr;
    WF_TRY(dalloc_t(h, "raw7", 7 * (size_t)n, &raw));
    for (int q = 0; q < 7; ++q) {
      WF_HIP(hipMemcpyAsync(raw + q * n, src[q], sizeof(float) * n, hipMemcpyHostToDevice, st));
      src[q] = raw + q * n;
    }
  }
  // validation + bounding box
  float* bb = nullptr;
  int* bad = nullptr;
  WF_TRY(dalloc_t(h, "bbox", 16, &bb));
  WF_TRY(dalloc_t(h, "bad", 4, &bad));
  WF_HIP(wf::launch_validate(src[0], src[1], src[2], src[3], src[4], src[5], src[6], n, bb, bad, st));
  float hb[6];
  int hbad = 0;
  WF_HIP(hipMemcpyAsync(hb, bb, sizeof hb, hipMemcpyDeviceToHost, st));
  WF_HIP(hipMemcpyAsync(&hbad, bad, sizeof hbad, hipMemcpyDeviceToHost, st));
  WF_HIP(hipStreamSynchronize(st));
  if (hbad)
    return set_err(WF_ERR_INVALID_ARG,
                   "surface: %d samples with non-finite values, |normal| far from 1 or area < 0", hbad);
  // root box (treePO.cu:220-230; automatic: bbox centre +- enlarge * half extent)
  if (h->opts.has_box) {
    for (int q = 0; q < 6; ++q) h->boxf[q] = (float)h->opts.box[q];
  } else {
    double half[3], cen[3], hmax = 0;
    for (int q = 0; q < 3; ++q) {
      cen[q] = 0.5 * ((double)hb[q] + (double)hb[q + 3]);
      half[q] = 0.5 * ((double)hb[q + 3] - (double)hb[q]) * h->opts.box_enlarge;
      hmax = std::max(hmax, half[q]);
    }
    if (hmax <= 0) hmax = 1e-9;
    for (int q = 0; q < 3; ++q) {
      if (half[q] < 1e-6 * hmax) half[q] = 1e-6 * hmax;
      h->boxf[q] = (float)(cen[q] - half[q]);
      h->boxf[q + 3] = (float)(cen[q] + half[q]);
    }
  }
  for (int q = 0; q < 6; ++q) h->box[q] = (double)h->boxf[q];
  // float split planes per axis and level (tree.h:116-120 centre, :672-687 boxes)
  const int per_axis = (1 << depth) - 1;
  std::vector<float> planes(3 * (size_t)std::max(per_axis, 1));
  for (int ax = 0; ax < 3; ++ax) {
    std::vector<float> lo(1, h->boxf[ax]), hi(1, h->boxf[ax + 3]);
    for (int l = 0; l < depth; ++l) {
      std::vector<float> nlo(2 << l), nhi(2 << l);
      for (int i = 0; i < (1 << l); ++i) {
        const float c = 0.5f * (lo[i] + hi[i]);
        planes[ax * per_axis + (1 << l) - 1 + i] = c;
        nlo[2 * i] = lo[i];
        nhi[2 * i] = c;
        nlo[2 * i + 1] = c;
        nhi[2 * i + 1] = hi[i];
      }
      lo.swap(nlo);
      hi.swap(nhi);
    }
  }
  float* planes_dev = nullptr;
  WF_TRY(dalloc_t(h, "planes", planes.size(), &planes_dev));
  WF_HIP(hipMemcpyAsync(planes_dev, planes.data(), sizeof(float) * planes.size(),
                        hipMemcpyHostToDevice, st));
  // keys + stable sort
  uint32_t *keys, *keys_tmp;
  int32_t *idx, *idx_tmp, *counts, *scan_tmp, *flags, *pos, *cnt_dev, *lpar;
  WF_TRY(dalloc_t(h, "keys", n, &keys));
  WF_TRY(dalloc_t(h, "keys_tmp", n, &keys_tmp));
  WF_TRY(dalloc_t(h, "perm", n, &idx));
  WF_TRY(dalloc_t(h, "idx_tmp", n, &idx_tmp));
  WF_TRY(dalloc_t(h, "sort_counts", wf::sort_count_ints(n) + 16, &counts));
  WF_TRY(dalloc_t(h, "scan_tmp", wf::scan_tmp_ints(std::max<int64_t>(n, wf::sort_count_ints(n))) + 16,
                  &scan_tmp));
  WF_TRY(dalloc_t(h, "flags", n, &flags));
  WF_TRY(dalloc_t(h, "pos", n, &pos));
  WF_TRY(dalloc_t(h, "level_counts", 16, &cnt_dev));
  WF_TRY(dalloc_t(h, "lvl_parent", (size_t)(depth + 1) * n, &lpar));
  WF_HIP(wf::octree_sort(src[0], src[1], src[2], n, planes_dev, depth, keys, idx, keys_tmp, idx_tmp,
                         counts, scan_tmp, st));
  float* pts = nullptr;
  WF_TRY(dalloc_t(h, "pts7", 7 * (size_t)n, &pts));
  WF_HIP(wf::gather_points(idx, n, src, pts, st));
  // per-level active lists
  h->levels.assign(depth + 1, LevelDev());
  uint32_t* lk = nullptr;
  int32_t *lb = nullptr, *le = nullptr, *lcb = nullptr, *lce = nullptr;
  WF_TRY(dalloc_t(h, "lvl_key", (size_t)(depth + 1) * n, &lk));
  WF_TRY(dalloc_t(h, "lvl_begin", (size_t)(depth + 1) * n, &lb));
  WF_TRY(dalloc_t(h, "lvl_end", (size_t)(depth + 1) * n, &le));
  WF_TRY(dalloc_t(h, "lvl_cb", (size_t)(depth + 1) * n, &lcb));
  WF_TRY(dalloc_t(h, "lvl_ce", (size_t)(depth + 1) * n, &lce));
  for (int l = 0; l <= depth; ++l) {
    LevelDev& L = h->levels[l];
    L.key = lk + (size_t)l * n;
    L.begin = lb + (size_t)l * n;
    L.end = le + (size_t)l * n;
    L.cb = lcb + (size_t)l * n;
    L.ce = lce + (size_t)l * n;
    L.parent = lpar + (size_t)l * n;
    WF_HIP(wf::octree_level(keys, n, 3 * (depth - l), flags, pos, scan_tmp, cnt_dev + l, L.key,
                            L.begin, st));
  }
  int hc[16] = {0};
  WF_HIP(hipMemcpyAsync(hc, cnt_dev, sizeof(int) * (depth + 1), hipMemcpyDeviceToHost, st));
  WF_HIP(hipStreamSynchronize(st));
  int32_t* spans = nullptr;
  WF_TRY(dalloc_t(h, "lvl_span", (size_t)2 * (depth + 1), &spans));
  std::vector<int32_t> hspan(2 * (depth + 1));
  for (int l = 0; l <= depth; ++l) {
    h->levels[l].count = hc[l];
    h->levels[l].span = spans + 2 * l;
    hspan[2 * l] = 0;
    hspan[2 * l + 1] = hc[l];
    WF_HIP(wf::octree_level_finish(h->levels[l].begin, hc[l], (int32_t)n, h->levels[l].end, st));
  }
  WF_HIP(hipMemcpyAsync(spans, hspan.data(), sizeof(int32_t) * hspan.size(), hipMemcpyHostToDevice,
                        st));
  WF_HIP(hipStreamSynchronize(st));
  for (int l = 0; l < depth; ++l)
    WF_HIP(wf::octree_links(h->levels[l + 1].begin, hc[l + 1], h->levels[l].begin, hc[l],
                            h->levels[l + 1].parent,
                            h->levels[l].cb, h->levels[l].ce, st));
  h->n = n;
  h->depth = depth;
  h->depth_max = depth;
  h->auto_depth = auto_depth;
  h->geom_gen++;
  h->pts = pts;
  h->perm = idx;
  h->keys = keys;
  h->has_geom = true;
  return WF_OK;
}

wf_status wf_octree_nodes(wf_handle* h, int32_t* perm, int32_t* node_begin, int32_t* node_end,
                          void* stream) {
  if (!h) return set_err(WF_ERR_INVALID_ARG, "handle is null");
  if (!h->has_geom) return set_err(WF_ERR_NO_GEOMETRY, "no geometry set");
  WF_TRY(set_device(h));
  hipStream_t st = as_stream(stream);
  if (perm)
    WF_HIP(hipMemcpyAsync(perm, h->perm, sizeof(int32_t) * h->n, hipMemcpyDeviceToDevice, st));
  if (node_begin && node_end)
    WF_HIP(wf::octree_node_ranges(h->keys, h->n, h->depth_max, node_begin, node_end, st));
  return WF_OK;
}

wf_status wf_currents(wf_handle* h, const wf_surface* s, const wf_incidence* inc, int32_t n_inc,
                      int32_t fp64, void* out, void* stream) {
  if (!h || !s || !out) return set_err(WF_ERR_INVALID_ARG, "null argument");
  if (s->memspace != WF_MEM_DEVICE) return set_err(WF_ERR_INVALID_ARG, "wf_currents needs device arrays");
  if (s->n <= 0) return set_err(WF_ERR_INVALID_ARG, "surface.n must be > 0");
  WF_TRY(check_inc(inc, n_inc));
  WF_TRY(set_device(h));
  hipStream_t st = as_stream(stream);
  WfIncDev* dinc = nullptr;
  WF_TRY(upload_incs(h, inc, n_inc, &dinc, st));
  WF_HIP(wf::launch_currents(s->x, s->y, s->z, s->nx, s->ny, s->nz, s->area, s->n, dinc, n_inc,
                             fp64 ? 1 : 0, out, 6, 6 * s->n, st));
  return WF_OK;
}

wf_status wf_interp(wf_handle* h, int32_t count, int32_t ntc, int32_t npc, const float* child,
                    int32_t ntp, int32_t npp, double off_tp, float* parent, void* stream) {
  if (!h || !child || !parent) return set_err(WF_ERR_INVALID_ARG, "null argument");
  if (count <= 0 || ntc < 1 || npc < 1 || ntp < 1 || npp < 2 || (npp % 2))
    return set_err(WF_ERR_INVALID_ARG, "wf_interp: bad sizes (npp must be even)");
  WF_TRY(set_device(h));
  hipStream_t st = as_stream(stream);
  FftPlanDev pc, pp, tc, tp;
  WF_TRY(get_fft(h, npc, &pc));
  WF_TRY(get_fft(h, npp, &pp));
  WF_TRY(get_fft(h, 2 * ntc, &tc));
  WF_TRY(get_fft(h, 2 * ntp, &tp));
  float2* G = nullptr;
  WF_TRY(dalloc_t(h, "interp_G",
                  std::max((size_t)count * wf::fft_g_stride(ntc) * npp,
                           (size_t)wf::gc_elems(count, ntc, npp)),
                  &G));
  const double off_tc = -WF_PI / 2 + WF_PI / (2.0 * ntc);
  const int mirror = (int)std::lround(ntp * (1.0 - 2.0 * off_tp / WF_PI));
  const bool mfma_ok = h->tune.agg_kernel == 0 && (mirror == 2 * ntp - 1 || mirror == 2 * ntp) &&
                       wf::gc_row_len(ntc) <= wf::kThetaMfmaMaxK && npc <= wf::kPhiMfmaMaxK &&
                       npc % 2 == 0 && npp % 2 == 0;
  const bool fft2_ok = h->tune.agg_kernel == 0 && (mirror == 2 * ntp - 1 || mirror == 2 * ntp) &&
                       is_smooth5(npc) && is_smooth5(npp) && is_smooth5(2 * ntc) &&
                       is_smooth5(2 * ntp);
  if (fft2_ok && (2 * ntp > h->tune.agg_fft_rows || !mfma_ok)) {
    Fft2Plan f_pc, f_pp, f_tc, f_tp;
    WF_TRY(get_fft2(h, npc, &f_pc));
    WF_TRY(get_fft2(h, npp, &f_pp));
    WF_TRY(get_fft2(h, 2 * ntc, &f_tc));
    WF_TRY(get_fft2(h, 2 * ntp, &f_tp));
    const float2 *fac_p = nullptr, *fac_t = nullptr;
    WF_TRY(get_fac(h, npc, 0.0, &fac_p));
    WF_TRY(get_fac(h, 2 * ntc, off_tp - off_tc, &fac_t));
    WF_HIP(wf::launch_phi_fft((const float2*)child, G, count, ntc, npc, npp, f_pc, f_pp, fac_p, st));
    wf::ThetaFftArgs ta;
    ta.GT = G;
    ta.F = 1;
    ta.ntc = ntc;
    ta.npp = npp;
    ta.ntp = ntp;
    ta.nparents = count;
    ta.cb = nullptr;
    ta.ce = nullptr;
    ta.child_key = nullptr;
    ta.Text = nullptr;
    ta.mirror = mirror;
    ta.pc = f_tc;
    ta.pp = f_tp;
    ta.fac = fac_t;
    ta.out = (float2*)parent;
    ta.out_root = nullptr;
    WF_HIP(wf::launch_theta_fft(ta, st));
    return WF_OK;
  }
  if (mfma_ok) {
    MatEntry mp, mt;
    WF_TRY(get_mat(h, npc, npp, 0.0, &mp));
    WF_TRY(get_mat(h, 2 * ntc, 2 * ntp, off_tp - off_tc, &mt, true));
    WF_HIP(wf::launch_phi_mfma((const float2*)child, G, (int64_t)count * ntc, ntc, npc, npp, mp.hi,
                               mp.lo, mp.kpad, st));
    wf::ThetaMfmaArgs ta;
    ta.GT = G;
    ta.F = 1;
    ta.ntc = ntc;
    ta.npp = npp;
    ta.ntp = ntp;
    ta.nparents = count;
    ta.cb = nullptr;
    ta.ce = nullptr;
    ta.child_key = nullptr;
    ta.Text = nullptr;
    ta.Mhi = mt.hi;
    ta.Mlo = mt.lo;
    ta.kpad = mt.kpad;
    ta.mirror = mirror;
    ta.out = (float2*)parent;
    ta.out_root = nullptr;
    WF_HIP(wf::launch_theta_mfma(ta, st));
    return WF_OK;
  }
  WF_HIP(wf::launch_phi_pass((const float2*)child, G, count, ntc, npc, npp, pc, pp, st));
  WF_HIP(wf::launch_theta_pass(G, 1, ntc, npp, ntp, count, nullptr, nullptr, nullptr, nullptr,
                               off_tp - off_tc, mirror, tc, tp, (float2*)parent, nullptr, st));
  return WF_OK;
}

static wf_status po_solve_fp64_impl(wf_handle* h, const wf_surface* s, const wf_incidence* inc,
                                  int32_t n_inc, wf_grid grid, double* intensity, double* E_cplx,
                                  void* stream, double dist) {
  if (!h) return set_err(WF_ERR_INVALID_ARG, "handle is null");
  if (!intensity) return set_err(WF_ERR_INVALID_ARG, "intensity is null");
  if (grid.n_theta < 1 || grid.n_phi < 1) return set_err(WF_ERR_INVALID_ARG, "grid must be >= 1x1");
  WF_TRY(check_inc(inc, n_inc));
  WF_TRY(wf_set_geometry(h, s, 0, stream));
  hipStream_t st = as_stream(stream);
  const int64_t D = (int64_t)grid.n_theta * grid.n_phi;
  const int64_t n = h->n;
  WfIncDev* dinc = nullptr;
  WF_TRY(upload_incs(h, inc, n_inc, &dinc, st));
  double* I_dev = intensity;
  double* E_dev = E_cplx;
  if (s->memspace == WF_MEM_HOST) {
    WF_TRY(dalloc_t(h, "out_Id", (size_t)n_inc * D, &I_dev));
    if (E_cplx) WF_TRY(dalloc_t(h, "out_Ed", (size_t)n_inc * 6 * D, &E_dev));
  }
  const double centre[3] = {0.5 * (h->box[0] + h->box[3]), 0.5 * (h->box[1] + h->box[4]),
                            0.5 * (h->box[2] + h->box[5])};
  const float* px = h->pts;
  // solves sharing a wavelength share the phase; at most 4 per launch (F <= 24)
  std::vector<int> order(n_inc);
  for (int i = 0; i < n_inc; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return inc[a].lambda_m < inc[b].lambda_m; });
  double2 *C = nullptr, *F0 = nullptr, *part = nullptr;
  WF_TRY(dalloc_t(h, "direct_C", (size_t)n * 24, &C));
  WF_TRY(dalloc_t(h, "direct_F0", (size_t)D * 24, &F0));
  const int64_t part_elems = (int64_t)wf::direct_split(D) * 24 * D;
  WF_TRY(dalloc_t(h, "direct_part", (size_t)part_elems, &part));
  prof_reset(h);
  WF_TRY(prof_event(h, st, &h->ev_t0));
  for (int i0 = 0; i0 < n_inc;) {
    int i1 = i0 + 1;
    while (i1 < n_inc && i1 - i0 < 4 && inc[order[i1]].lambda_m == inc[order[i0]].lambda_m) ++i1;
    const int S = i1 - i0;
    if (cancelled(h)) return set_err(WF_ERR_CANCELLED, "cancelled");
    for (int q = 0; q < S; ++q)
      WF_TRY(timed(h, st, WF_K_CURRENTS, (28.0 + 96.0) * n, [&] {
        return wf::launch_currents(px, px + n, px + 2 * n, px + 3 * n, px + 4 * n, px + 5 * n,
                                   px + 6 * n, n, dinc + order[i0 + q], 1, 1, C + 6 * q, 6 * S, 0, st);
      }));
    WF_TRY(timed(h, st, WF_K_DIRECT, 48.0 * n * D * S, [&] {
      return wf::launch_direct_fp64(px, px + n, px + 2 * n, n, C, 6 * S, centre,
                                    2.0 * WF_PI / inc[order[i0]].lambda_m, grid.n_theta,
                                    grid.n_phi, F0, part, part_elems, st, dist);
    }));
    for (int q = 0; q < S; ++q) {
      const int slot = order[i0 + q];
      WF_TRY(timed(h, st, WF_K_EPILOGUE, (96.0 + 8.0 + (E_dev ? 48.0 : 0.0)) * D, [&] {
        return wf::launch_intensity(F0 + (int64_t)q * 6 * D, 1, grid.n_theta, grid.n_phi, nullptr,
                                    nullptr, I_dev + (int64_t)slot * D,
                                    E_dev ? E_dev + (int64_t)slot * 6 * D : nullptr, st);
      }));
    }
    i0 = i1;
  }
  WF_TRY(prof_event(h, st, &h->ev_t1));
  if (s->memspace == WF_MEM_HOST) {
    WF_HIP(hipMemcpyAsync(intensity, I_dev, sizeof(double) * n_inc * D, hipMemcpyDeviceToHost, st));
    if (E_cplx)
      WF_HIP(hipMemcpyAsync(E_cplx, E_dev, sizeof(double) * n_inc * 6 * D, hipMemcpyDeviceToHost, st));
    WF_HIP(hipStreamSynchronize(st));
  }
  return WF_OK;
}

wf_status wf_po_solve_fp64_direct(wf_handle* h, const wf_surface* s, const wf_incidence* inc,
                                  int32_t n_inc, wf_grid grid, double* intensity, double* E_cplx,
                                  void* stream) {
  return po_solve_fp64_impl(h, s, inc, n_inc, grid, intensity, E_cplx, stream, 0.0);
}

wf_status wf_po_solve_fp64_finite(wf_handle* h, const wf_surface* s, const wf_incidence* inc,
                                  int32_t n_inc, wf_grid grid, double distance_m,
                                  double* intensity, double* E_cplx, void* stream) {
  if (!(distance_m > 0 && std::isfinite(distance_m)))
    return set_err(WF_ERR_INVALID_ARG, "distance must be positive and finite");
  return po_solve_fp64_impl(h, s, inc, n_inc, grid, intensity, E_cplx, stream, distance_m);
}

}  // extern "C"

namespace {

// Per-wavelength tree context: level grids and translation tables.
struct GroupCtx {
  double lam = 0, k0 = 0;
  std::vector<Grid> g;
  float2* T = nullptr;
  std::vector<int64_t> toff;
  float2* Tx = nullptr;  // extended tables of the MFMA level transition
  std::vector<int64_t> txoff;
  int root_from = 1;         // level whose patterns the root evaluates
  float2* Troot = nullptr;   // per-child table of that transition (root_from > 1)
};

// Which level-transition kernels can take child grid gcg -> parent grid gp:
// the dense MFMA pair (phi_tile + theta_mfma) and/or the LDS FFT pair; with
// neither, the older FFT kernels run (they read the plain translation tables).
struct LevelPath {
  bool mfma, fft2;
};
LevelPath level_path(const wf_handle* h, Grid gcg, Grid gp);

double theta_origin(int l, int nt) {
  return (l == 0) ? (-WF_PI / 2 + 1e-10) : (-WF_PI / 2 + WF_PI / (2.0 * nt));
}

int mirror_of(int nt, double off) { return (int)std::lround(nt * (1.0 - 2.0 * off / WF_PI)); }

// Output grids of a tree solve: even n_phi (theta extension), FFT lengths.
wf_status check_tree_grid(wf_grid grid) {
  const int NT = grid.n_theta, NP = grid.n_phi;
  if (NP % 2) return set_err(WF_ERR_INVALID_ARG, "tree solve needs an even n_phi (theta extension)");
  if (!is_smooth(NP) || !is_smooth(2 * NT))
    return set_err(WF_ERR_UNSUPPORTED, "output grid %dx%d: n_phi and 2 n_theta must be 13-smooth",
                   NT, NP);
  return WF_OK;
}

double sample_lambda(const wf_handle* h, double lam) {
  return h->opts.lambda_sample_m > 0 ? h->opts.lambda_sample_m : lam;
}

// Modelled seconds per solve of a depth-d tree pass (see wf_tune).
double model_cost(const wf_handle* h, const wf_tune& t, int d, double lam_s, int NT, int NP) {
  std::vector<Grid> g(d + 1, Grid{NT, NP});
  for (int l = 1; l <= d; ++l) g[l] = level_grid(h, l, lam_s);
  double bytes = 0;
  for (int l = 0; l < d; ++l)
    bytes += 48.0 * ((double)h->levels[l + 1].count * g[l + 1].nt * g[l + 1].np +
                     (double)h->levels[l].count * g[l].nt * g[l].np);
  return t.leaf_s_per_pair * (double)h->n * g[d].nt * g[d].np + t.agg_s_per_byte * bytes +
         t.level_s * d;
}

// Depth of a wavelength group: the built depth, or for WF_DEPTH_AUTO
// geometries the model's cheapest depth in [1, depth_max].
int choose_depth(const wf_handle* h, double lam, int NT, int NP) {
  if (h->force_depth >= 0) return h->force_depth;
  if (!h->auto_depth) return h->depth_max;
  const double lam_s = sample_lambda(h, lam);
  if (h->tuned_depth >= 0 && h->tuned_gen == h->geom_gen && h->tuned_lam == lam_s)
    return std::min(h->tuned_depth, h->depth_max);
  int best = h->depth_max;
  double cbest = 0;
  for (int d = 1; d <= h->depth_max; ++d) {
    const double c = model_cost(h, h->tune, d, lam_s, NT, NP);
    if (d == 1 || c < cbest) {
      cbest = c;
      best = d;
    }
  }
  return best;
}

// Plain translation tables [level][8 octants][nt][np] (read by the older FFT
// kernels and by the adjoint; the MFMA and LDS FFT transitions use the
// extended tables).
wf_status build_plain_tables(wf_handle* h, hipStream_t st, GroupCtx& gc, int depth) {
  const int64_t tt = gc.toff[depth - 1] + 8LL * gc.g[depth - 1].nt * gc.g[depth - 1].np;
  WF_TRY(dalloc_t(h, "trans", (size_t)tt, &gc.T));
  for (int l = 0; l < depth; ++l) {
    const double th0 =
        (l == 0) ? (-WF_PI / 2 + 1e-10) : (-WF_PI / 2 + WF_PI / (2.0 * gc.g[l].nt));
    const double hc[3] = {0.5 * (h->box[3] - h->box[0]) / std::ldexp(1.0, l + 1),
                          0.5 * (h->box[4] - h->box[1]) / std::ldexp(1.0, l + 1),
                          0.5 * (h->box[5] - h->box[2]) / std::ldexp(1.0, l + 1)};
    const Grid gl = gc.g[l];
    float2* Tl = gc.T + gc.toff[l];
    const double k0 = gc.k0;
    WF_TRY(timed(h, st, WF_K_TRANS, 64.0 * gl.nt * gl.np,
                 [&] { return wf::launch_trans_table(Tl, gl.nt, gl.np, th0, hc, k0, st); }));
  }
  return WF_OK;
}

LevelPath level_path(const wf_handle* h, Grid gcg, Grid gp) {
  LevelPath p;
  p.mfma = h->tune.agg_kernel == 0 && wf::gc_row_len(gcg.nt) <= wf::kThetaMfmaMaxK &&
           gcg.np <= wf::kPhiMfmaMaxK && gcg.np % 2 == 0 && gp.np % 2 == 0;
  p.fft2 = h->tune.agg_kernel == 0 && is_smooth5(gp.np) && is_smooth5(2 * gp.nt) &&
           is_smooth5(gcg.np) && is_smooth5(2 * gcg.nt);
  return p;
}

// The transition child grid -> parent grid runs the LDS FFT kernels (long
// theta rows, or grids the MFMA kernels do not take).
bool fft_transition(const wf_handle* h, Grid gcg, Grid gp) {
  const LevelPath lp = level_path(h, gcg, gp);
  return lp.fft2 && (2 * gp.nt > h->tune.agg_fft_rows || !lp.mfma);
}

// Translation-table layout of a transition: k-major for the wave-private LDS
// FFT theta kernel.
int table_kmajor(const wf_handle* h, Grid gcg, Grid gp) {
  return (fft_transition(h, gcg, gp) && wf::theta_fft_wave(gcg.nt, gp.nt)) ? 1 : 0;
}

// Grids and translation tables of one wavelength group.  allow_skip: the
// root may read the first level whose grid is no larger than the output grid
// directly (forward solves; the adjoint keeps every level).
wf_status prepare_group(wf_handle* h, hipStream_t st, double lam, int NT, int NP, GroupCtx& gc,
                        bool allow_skip = true) {
  const int depth = h->depth;
  gc.lam = lam;
  gc.k0 = 2.0 * WF_PI / lam;
  const double lam_s = sample_lambda(h, lam);
  gc.g.assign(depth + 1, Grid{NT, NP});
  gc.toff.assign(depth + 1, 0);
  if (depth == 0) return WF_OK;
  for (int l = 1; l <= depth; ++l) gc.g[l] = level_grid(h, l, lam_s);
  if (h->opts.strict_grid && (NT < gc.g[1].nt || NP < gc.g[1].np))
    return set_err(WF_ERR_GRID_TOO_COARSE, "output grid %dx%d coarser than level-1 grid %dx%d",
                   NT, NP, gc.g[1].nt, gc.g[1].np);
  int64_t tt = 0;
  for (int l = 0; l < depth; ++l) {
    gc.toff[l] = tt;
    tt += 8LL * gc.g[l].nt * gc.g[l].np;
  }
  gc.T = nullptr;  // built below only when a kernel that reads it will run
  if (h->tune.agg_kernel == 0) {
    gc.txoff.assign(depth + 1, 0);
    int64_t tx = 0;
    for (int l = 0; l < depth; ++l) {
      gc.txoff[l] = tx;
      tx += 8LL * (2 * gc.g[l].nt + 1) * (gc.g[l].np / 2);
    }
    WF_TRY(dalloc_t(h, "trans_ext", (size_t)tx, &gc.Tx));
    for (int l = 0; l < depth; ++l) {
      const Grid gl = gc.g[l];
      const double th0 = theta_origin(l, gl.nt);
      const int mirror = mirror_of(gl.nt, th0);
      const double hc[3] = {0.5 * (h->box[3] - h->box[0]) / std::ldexp(1.0, l + 1),
                            0.5 * (h->box[4] - h->box[1]) / std::ldexp(1.0, l + 1),
                            0.5 * (h->box[5] - h->box[2]) / std::ldexp(1.0, l + 1)};
      float2* Tl = gc.Tx + gc.txoff[l];
      const double k0 = gc.k0;
      const int kmajor = table_kmajor(h, gc.g[l + 1], gl);
      WF_TRY(timed(h, st, WF_K_TRANS, 8.0 * 8 * (2 * gl.nt + 1) * (gl.np / 2), [&] {
        return wf::launch_trans_ext(Tl, gl.nt, gl.np, mirror, th0, hc, k0, st, kmajor);
      }));
    }
  }
  gc.root_from = 1;
  gc.Troot = nullptr;
  if (allow_skip && h->tune.agg_kernel == 0 && depth >= 2) {
    int r = 1;
    while (r < depth && (gc.g[r].nt > NT || gc.g[r].np > NP)) ++r;
    const Grid gr = gc.g[r];
    const double off_p = theta_origin(0, NT);
    const int mirror = mirror_of(NT, off_p);
    const bool mfma_ok = wf::gc_row_len(gr.nt) <= wf::kThetaMfmaMaxK &&
                         gr.np <= wf::kPhiMfmaMaxK && gr.np % 2 == 0 && NP % 2 == 0;
    const bool fft2_ok = is_smooth5(NP) && is_smooth5(2 * NT) && is_smooth5(gr.np) &&
                         is_smooth5(2 * gr.nt);
    const int nr = h->levels[r].count;
    const double tbytes = 8.0 * nr * (2.0 * NT + 1) * (NP / 2);
    if (r > 1 && (mfma_ok || fft2_ok) && (mirror == 2 * NT - 1 || mirror == 2 * NT) &&
        tbytes <= 2.0 * 1024 * 1024 * 1024) {
      double hx[16], hy[16], hz[16];
      for (int l = 0; l < r; ++l) {
        hx[l] = 0.5 * (h->box[3] - h->box[0]) / std::ldexp(1.0, l + 1);
        hy[l] = 0.5 * (h->box[4] - h->box[1]) / std::ldexp(1.0, l + 1);
        hz[l] = 0.5 * (h->box[5] - h->box[2]) / std::ldexp(1.0, l + 1);
      }
      double* coff = nullptr;
      WF_TRY(dalloc_t(h, "root_offsets", (size_t)3 * nr, &coff));
      WF_HIP(wf::launch_child_offsets(h->levels[r].key, nr, r, hx, hy, hz, coff, coff + nr,
                                      coff + 2 * nr, st));
      WF_TRY(dalloc_t(h, "trans_root", (size_t)nr * (2 * NT + 1) * (NP / 2), &gc.Troot));
      const double k0 = gc.k0;
      const int kmajor = table_kmajor(h, gr, gc.g[0]);
      WF_TRY(timed(h, st, WF_K_TRANS, tbytes, [&] {
        return wf::launch_trans_child(gc.Troot, NT, NP, mirror, off_p, coff, coff + nr,
                                      coff + 2 * nr, nr, k0, st, kmajor);
      }));
      gc.root_from = r;
    }
  }
  // the plain tables: for the adjoint, and for transitions that neither the
  // MFMA nor the LDS FFT kernels take (the dispatch of forward_chunk)
  bool need_plain = !allow_skip;
  for (int l = depth - 1; l >= 0 && !need_plain; --l) {
    if (l > 0 && l < gc.root_from) continue;
    const LevelPath lp = level_path(h, gc.g[(l == 0) ? gc.root_from : l + 1], gc.g[l]);
    if (!lp.mfma && !lp.fft2) need_plain = true;
  }
  if (need_plain) WF_TRY(build_plain_tables(h, st, gc, depth));
  for (int l = 0; l <= depth && l < 16; ++l) {
    h->stats.level_nt[l] = gc.g[l].nt;
    h->stats.level_np[l] = gc.g[l].np;
    h->stats.level_nodes[l] = h->levels[l].count;
  }
  h->stats.n_levels = depth + 1;
  return WF_OK;
}

wf::LeafArgs leaf_args(wf_handle* h, const GroupCtx& gc, const float2* C, int F, float2* out) {
  const int depth = h->depth;
  const int64_t n = h->n;
  const LevelDev& LL = h->levels[depth];
  wf::LeafArgs la;
  la.px = h->pts;
  la.py = h->pts + n;
  la.pz = h->pts + 2 * n;
  la.n = n;
  la.C = C;
  la.Ftot = F;
  la.f0 = 0;
  la.F = F;
  la.leaf_begin = LL.begin;
  la.leaf_end = LL.end;
  la.leaf_key = LL.key;
  la.nleaves = LL.count;
  la.depth = depth;
  for (int q = 0; q < 3; ++q) {
    la.box_lo[q] = h->box[q];
    la.leaf_size[q] = (h->box[q + 3] - h->box[q]) / std::ldexp(1.0, depth);
  }
  la.k0 = gc.k0;
  la.nt = gc.g[depth].nt;
  la.np = gc.g[depth].np;
  la.out = out;
  la.precision = h->opts.precision;
  la.quad = h->tune.leaf_quad;
  return la;
}

struct TreeBufs {
  float2 *patA = nullptr, *patB = nullptr, *G = nullptr, *G0 = nullptr;
};

wf_status tree_buffers(wf_handle* h, const GroupCtx& gc, int F, int NP, TreeBufs& tb) {
  const int depth = h->depth;
  size_t pat_max = 16, g_max = 16;
  for (int l = 1; l <= depth; ++l)
    pat_max = std::max(pat_max, (size_t)h->levels[l].count * F * gc.g[l].nt * gc.g[l].np);
  // G holds the phi-pass output of one level: [child][F][np_parent][fft_g_stride(nt_child)]
  // for the FFT kernels, padded great-circle rows for the MFMA kernels
  auto g_elems = [&](int64_t nchild, int ntc, int npp) {
    return std::max((size_t)nchild * F * wf::fft_g_stride(ntc) * npp,
                    (size_t)wf::gc_elems(nchild * F, ntc, npp));
  };
  for (int l = 1; l < depth; ++l)
    g_max = std::max(g_max, g_elems(h->levels[l + 1].count, gc.g[l + 1].nt, gc.g[l].np));
  WF_TRY(dalloc_t(h, "patA", pat_max, &tb.patA));
  WF_TRY(dalloc_t(h, "patB", pat_max, &tb.patB));
  WF_TRY(dalloc_t(h, "G", g_max, &tb.G));
  const int r = gc.root_from;
  WF_TRY(dalloc_t(h, "G0", g_elems(h->levels[r].count, gc.g[r].nt, NP), &tb.G0));
  return WF_OK;
}

// Forward tree pass of one chunk of S solves sharing a wavelength: currents ->
// leaf sums -> aggregation -> root evaluation on the output grid.  F0 receives
// the FP64 far-field components [6 S][Nt*Np].
// transverse: aggregate the two transverse far-field components of each
// solve instead of its six fields (wf_transverse.hip) when every transition
// of the chunk runs a kernel that carries the great circle's sign flips;
// *fps receives the fields per solve of F0 (2 or 6).
wf_status forward_chunk(wf_handle* h, hipStream_t st, const GroupCtx& gc, const WfIncDev* dinc,
                        const int* order, int c0, int S, int NT, int NP, const TreeBufs& tb,
                        double2* F0, bool transverse, int* fps) {
  const int depth = h->depth;
  const int64_t n = h->n;
  const int64_t D = (int64_t)NT * NP;
  const int F6 = 6 * S;
  const float* px = h->pts;
  float2* C = nullptr;
  WF_TRY(dalloc_t(h, "C", (size_t)n * F6, &C));
  // one launch when the chunk's incidences are consecutive in dinc
  bool consecutive = true;
  for (int q = 1; q < S; ++q) consecutive = consecutive && order[c0 + q] == order[c0] + q;
  if (consecutive) {
    WF_TRY(timed(h, st, WF_K_CURRENTS, (28.0 + 48.0) * n * S, [&] {
      return wf::launch_currents(px, px + n, px + 2 * n, px + 3 * n, px + 4 * n, px + 5 * n,
                                 px + 6 * n, n, dinc + order[c0], S, 0, C, F6, 6, st);
    }));
  } else {
    for (int q = 0; q < S; ++q)
      WF_TRY(timed(h, st, WF_K_CURRENTS, (28.0 + 48.0) * n, [&] {
        return wf::launch_currents(px, px + n, px + 2 * n, px + 3 * n, px + 4 * n, px + 5 * n,
                                   px + 6 * n, n, dinc + order[c0 + q], 1, 0, C + 6 * q, F6, 0,
                                   st);
      }));
  }
  // the root reads level root_from directly when the levels above it have
  // grids larger than the output grid (their transitions would only cost)
  const int r = gc.root_from;
  if (transverse) {
    for (int l = depth - 1; l >= 0 && transverse; --l) {
      if (l > 0 && l < r) continue;
      const LevelPath lp = level_path(h, gc.g[(l == 0) ? r : l + 1], gc.g[l]);
      if (!lp.mfma && !lp.fft2) transverse = false;
    }
  }
  const int F = transverse ? 2 * S : F6;  // fields per level pattern
  *fps = transverse ? 2 : 6;
  float2* cur = (depth % 2) ? tb.patA : tb.patB;
  wf::LeafArgs la = leaf_args(h, gc, C, F6, cur);
  const double dleaf = (double)gc.g[depth].nt * gc.g[depth].np;
  const double z0 = std::sqrt((4e-7 * WF_PI) / 8.8541878e-12);  // as the epilogue
  const bool bf16 = h->opts.precision == WF_PREC_BF16X3;
  // the quad-symmetric leaf kernel projects onto the transverse components in
  // its epilogue; other leaf kernels write the six fields for the projection
  // kernel below
  const bool leaf_proj = transverse && h->opts.precision != WF_PREC_FP32_VALU &&
                         wf::leaf_fuses_transverse(la, bf16);
  la.transverse = leaf_proj ? 1 : 0;
  la.z0 = (float)z0;
  WF_TRY(timed(h, st, WF_K_LEAF, 8.0 * dleaf * n * F6, [&] {
    if (h->opts.precision == WF_PREC_FP32_VALU) return wf::launch_leaf(la, st);
    return wf::launch_leaf_mfma(la, bf16, st);
  }));
  h->stats.leaf_pairs += (int64_t)(dleaf * n);
  if (transverse && !leaf_proj) {
    float2* nxt = (cur == tb.patA) ? tb.patB : tb.patA;
    const int nleaf = h->levels[depth].count;
    WF_TRY(timed(h, st, WF_K_TRANS, 8.0 * (F6 + F) * nleaf * dleaf, [&] {
      return wf::launch_project_transverse(cur, nxt, nleaf, S, gc.g[depth].nt, gc.g[depth].np,
                                           theta_origin(depth, gc.g[depth].nt), z0, st);
    }));
    cur = nxt;
  }
  for (int l = depth - 1; l >= 0; --l) {
    if (l > 0 && l < r) continue;
    if (cancelled(h)) return set_err(WF_ERR_CANCELLED, "cancelled");
    const int lc = (l == 0) ? r : l + 1;  // child level
    const bool direct_root = (l == 0 && r > 1);
    const int32_t* pcb = direct_root ? h->levels[lc].span : h->levels[l].cb;
    const int32_t* pce = direct_root ? h->levels[lc].span + 1 : h->levels[l].ce;
    // extended tables exist only for the MFMA / LDS FFT kernels (agg_kernel 0)
    const float2* Tx_l = direct_root ? gc.Troot
                                     : (gc.txoff.empty() ? nullptr : gc.Tx + gc.txoff[l]);
    const Grid gcg = gc.g[lc], gp = gc.g[l];
    FftPlanDev pc, pp, tc, tp;
    WF_TRY(get_fft(h, gcg.np, &pc));
    WF_TRY(get_fft(h, gp.np, &pp));
    WF_TRY(get_fft(h, 2 * gcg.nt, &tc));
    WF_TRY(get_fft(h, 2 * gp.nt, &tp));
    const int nchild = h->levels[lc].count, npar = h->levels[l].count;
    const double planes = (double)nchild * F;
    float2* G = (l == 0) ? tb.G0 : tb.G;
    const double off_c = theta_origin(lc, gcg.nt), off_p = theta_origin(l, gp.nt);
    const int mirror = mirror_of(gp.nt, off_p);
    const float2* Tl = gc.T ? gc.T + gc.toff[l] : nullptr;
    const LevelPath lpath = level_path(h, gcg, gp);
    const bool mfma_ok = lpath.mfma;
    // compulsory bytes of this transition: children read once, parents
    // written once (the FP64 root accumulators at 16 B per sample)
    h->stats.agg_transitions += 1;
    h->stats.agg_bytes += 8.0 * planes * gcg.nt * gcg.np +
                          ((l == 0) ? 16.0 * F * D : 8.0 * F * (double)npar * gp.nt * gp.np);
    if (fft_transition(h, gcg, gp)) {
      // large grids: LDS FFT (each sample read and written once)
      Fft2Plan f_pc, f_pp, f_tc, f_tp;
      WF_TRY(get_fft2(h, gcg.np, &f_pc));
      WF_TRY(get_fft2(h, gp.np, &f_pp));
      WF_TRY(get_fft2(h, 2 * gcg.nt, &f_tc));
      WF_TRY(get_fft2(h, 2 * gp.nt, &f_tp));
      const float2 *fac_p = nullptr, *fac_t = nullptr;
      WF_TRY(get_fac(h, gcg.np, 0.0, &fac_p));
      WF_TRY(get_fac(h, 2 * gcg.nt, off_p - off_c, &fac_t));
      WF_TRY(timed(h, st, WF_K_PHI, 8.0 * planes * gcg.nt * ((double)gcg.np + gp.np), [&] {
        return wf::launch_phi_fft(cur, G, (int64_t)nchild * F, gcg.nt, gcg.np, gp.np, f_pc, f_pp,
                                  fac_p, st);
      }));
      float2* nxt = (cur == tb.patA) ? tb.patB : tb.patA;
      wf::ThetaFftArgs ta;
      ta.GT = G;
      ta.F = F;
      ta.ntc = gcg.nt;
      ta.npp = gp.np;
      ta.ntp = gp.nt;
      ta.nparents = npar;
      ta.cb = pcb;
      ta.ce = pce;
      ta.child_key = h->levels[lc].key;
      ta.Text = Tx_l;
      ta.text_per_child = direct_root ? 1 : 0;
      ta.mirror = mirror;
      ta.pc = f_tc;
      ta.pp = f_tp;
      ta.fac = fac_t;
      ta.out = (l == 0) ? nullptr : nxt;
      ta.out_root = (l == 0) ? F0 : nullptr;
      ta.flip = transverse ? 1 : 0;
      ta.text_kmajor = table_kmajor(h, gcg, gp);
      const double wbytes = (l == 0) ? 16.0 * F * D : 8.0 * F * (double)npar * gp.nt * gp.np;
      WF_TRY(timed(h, st, WF_K_THETA, 8.0 * planes * gcg.nt * gp.np + wbytes,
                   [&] { return wf::launch_theta_fft(ta, st); }));
      if (l > 0) cur = nxt;
      continue;
    }
    if (mfma_ok && l > 0 && h->tune.agg_fused && Tx_l &&
        wf::agg_fused_fits(gcg.nt, gcg.np, gp.nt, gp.np) && mirror == 2 * gp.nt - 1) {
      // one kernel: phi GEMM into LDS, theta GEMM, translation, child sum
      MatEntry mp, mt;
      WF_TRY(get_mat(h, gcg.np, gp.np, 0.0, &mp));
      WF_TRY(get_mat(h, 2 * gcg.nt, 2 * gp.nt, off_p - off_c, &mt, true, transverse));
      float2* nxt = (cur == tb.patA) ? tb.patB : tb.patA;
      wf::AggFusedArgs fa;
      fa.X = cur;
      fa.out = nxt;
      fa.F = F;
      fa.ntc = gcg.nt;
      fa.npc = gcg.np;
      fa.ntp = gp.nt;
      fa.npp = gp.np;
      fa.nparents = npar;
      fa.cb = pcb;
      fa.ce = pce;
      fa.child_key = h->levels[lc].key;
      fa.Text = Tx_l;
      fa.Mphi_hi = mp.hi;
      fa.Mphi_lo = mp.lo;
      fa.kpad_phi = mp.kpad;
      fa.mrows_phi = mp.rows_pad;
      fa.Mth_hi = mt.hi;
      fa.Mth_lo = mt.lo;
      fa.kpad_th = mt.kpad;
      fa.mrows_th = mt.rows_pad;
      fa.mirror = mirror;
      fa.version = h->tune.agg_fused;
      const double cbytes = 8.0 * planes * gcg.nt * gcg.np;
      const double pbytes = 8.0 * F * (double)npar * gp.nt * gp.np;
      WF_TRY(timed(h, st, WF_K_THETA, cbytes + pbytes,
                   [&] { return wf::launch_agg_fused_mfma(fa, st); }));
      cur = nxt;
      continue;
    }
    if (mfma_ok) {
      // dense split-BF16 MFMA interpolation: phi GEMM -> GT, theta GEMM +
      // translation + child sum -> parent (FP64 at the root)
      MatEntry mp, mt;
      WF_TRY(get_mat(h, gcg.np, gp.np, 0.0, &mp));
      WF_TRY(get_mat(h, 2 * gcg.nt, 2 * gp.nt, off_p - off_c, &mt, true, transverse));
      const int64_t rows = (int64_t)nchild * F * gcg.nt;
      WF_TRY(timed(h, st, WF_K_PHI, 8.0 * planes * gcg.nt * ((double)gcg.np + gp.np), [&] {
        return wf::launch_phi_mfma(cur, G, rows, gcg.nt, gcg.np, gp.np, mp.hi, mp.lo, mp.kpad, st,
                                   h->tune.phi_rows);
      }));
      float2* nxt = (cur == tb.patA) ? tb.patB : tb.patA;
      wf::ThetaMfmaArgs ta;
      ta.GT = G;
      ta.F = F;
      ta.ntc = gcg.nt;
      ta.npp = gp.np;
      ta.ntp = gp.nt;
      ta.nparents = npar;
      ta.cb = pcb;
      ta.ce = pce;
      ta.child_key = h->levels[lc].key;
      ta.Text = Tx_l;
      ta.text_per_child = direct_root ? 1 : 0;
      ta.Mhi = mt.hi;
      ta.Mlo = mt.lo;
      ta.kpad = mt.kpad;
      ta.mirror = mirror;
      ta.out = (l == 0) ? nullptr : nxt;
      ta.out_root = (l == 0) ? F0 : nullptr;
      ta.tile = h->tune.theta_tile;
      ta.flip = transverse ? 1 : 0;
      const double wbytes = (l == 0) ? 16.0 * F * D : 8.0 * F * (double)npar * gp.nt * gp.np;
      WF_TRY(timed(h, st, WF_K_THETA, 8.0 * planes * gcg.nt * gp.np + wbytes,
                   [&] { return wf::launch_theta_mfma(ta, st); }));
      if (l > 0) cur = nxt;
      continue;
    }
    if (l > 0 && wf::fused_fits(gcg.nt, gcg.np, gp.nt, gp.np)) {
      // one kernel: child read once, parent written once
      float2* nxt = (cur == tb.patA) ? tb.patB : tb.patA;
      WF_TRY(timed(h, st, WF_K_THETA,
                   8.0 * F * ((double)nchild * gcg.nt * gcg.np + (double)npar * gp.nt * gp.np), [&] {
        return wf::launch_agg_fused(cur, nxt, F, gcg.nt, gcg.np, gp.nt, gp.np, npar,
                                    h->levels[l].cb, h->levels[l].ce, h->levels[l + 1].key, Tl,
                                    off_p - off_c, mirror, pc, pp, tc, tp, st);
      }));
      cur = nxt;
      continue;
    }
    WF_TRY(timed(h, st, WF_K_PHI, 8.0 * planes * gcg.nt * ((double)gcg.np + gp.np), [&] {
      return wf::launch_phi_pass(cur, G, (int64_t)nchild * F, gcg.nt, gcg.np, gp.np, pc, pp, st);
    }));
    if (l == 0) {
      WF_TRY(timed(h, st, WF_K_THETA, 8.0 * planes * gcg.nt * NP + 16.0 * F * D, [&] {
        return wf::launch_theta_pass(G, F, gcg.nt, NP, NT, 1, h->levels[0].cb, h->levels[0].ce,
                                     h->levels[1].key, Tl, off_p - off_c, mirror, tc, tp, nullptr,
                                     F0, st);
      }));
    } else {
      float2* nxt = (cur == tb.patA) ? tb.patB : tb.patA;
      WF_TRY(timed(h, st, WF_K_THETA,
                   8.0 * F * ((double)nchild * gcg.nt * gp.np + (double)npar * gp.nt * gp.np), [&] {
        return wf::launch_theta_pass(G, F, gcg.nt, gp.np, gp.nt, npar, h->levels[l].cb,
                                     h->levels[l].ce, h->levels[l + 1].key, Tl, off_p - off_c,
                                     mirror, tc, tp, nxt, nullptr, st);
      }));
      cur = nxt;
    }
  }
  return WF_OK;
}

// Reverse tree pass of one chunk: root cogradient gF [F][D] (FP64) -> leaf
// cogradients -> per-point back-projection -> gradient accumulation (acc).
wf_status backward_chunk(wf_handle* h, hipStream_t st, const GroupCtx& gc, const WfIncDev* dinc,
                         const int* order, int c0, int S, int NT, int NP, const TreeBufs& tb,
                         const double2* gF, double* acc) {
  const int depth = h->depth;
  const int64_t n = h->n;
  const int F = 6 * S;
  float2* gcur = nullptr;  // cogradient of the current level's patterns
  for (int l = 0; l < depth; ++l) {
    const Grid gp = gc.g[l], gch = gc.g[l + 1];
    FftPlanDev pc, pp, tc, tp;
    WF_TRY(get_fft(h, gch.np, &pc));
    WF_TRY(get_fft(h, gp.np, &pp));
    WF_TRY(get_fft(h, 2 * gch.nt, &tc));
    WF_TRY(get_fft(h, 2 * gp.nt, &tp));
    const int nchild = h->levels[l + 1].count;
    const double off_c = theta_origin(l + 1, gch.nt), off_p = theta_origin(l, gp.nt);
    const int mirror = mirror_of(gp.nt, off_p);
    float2* G = (l == 0) ? tb.G0 : tb.G;
    const float2* Tl = gc.T + gc.toff[l];
    WF_TRY(timed(h, st, WF_K_THETA, 0.0, [&] {
      return wf::launch_theta_adj(l == 0 ? nullptr : gcur, l == 0 ? gF : nullptr, F, gch.nt,
                                  gp.np, gp.nt, nchild, h->levels[l + 1].parent,
                                  h->levels[l + 1].key, Tl, off_p - off_c, mirror, tp, tc, G, st);
    }));
    float2* nxt = (gcur == tb.patA) ? tb.patB : tb.patA;
    WF_TRY(timed(h, st, WF_K_PHI, 0.0, [&] {
      return wf::launch_phi_adj(G, nxt, (int64_t)nchild * F, gch.nt, gch.np, gp.np, pp, pc, st);
    }));
    gcur = nxt;
  }
  float2* Q = nullptr;
  WF_TRY(dalloc_t(h, "adj_Q", (size_t)n * F * 4, &Q));
  wf::LeafAdjArgs aa;
  const LevelDev& LL = h->levels[depth];
  aa.px = h->pts;
  aa.py = h->pts + n;
  aa.pz = h->pts + 2 * n;
  aa.leaf_begin = LL.begin;
  aa.leaf_end = LL.end;
  aa.leaf_key = LL.key;
  aa.nleaves = LL.count;
  aa.depth = depth;
  for (int q = 0; q < 3; ++q) {
    aa.box_lo[q] = h->box[q];
    aa.leaf_size[q] = (h->box[q + 3] - h->box[q]) / std::ldexp(1.0, depth);
  }
  aa.k0 = gc.k0;
  aa.nt = gc.g[depth].nt;
  aa.np = gc.g[depth].np;
  aa.F = F;
  aa.g = gcur;
  aa.Q = Q;
  WF_TRY(timed(h, st, WF_K_LEAF, 0.0, [&] { return wf::launch_leaf_adj(aa, st); }));
  for (int q = 0; q < S; ++q)
    WF_TRY(timed(h, st, WF_K_CURRENTS, 0.0, [&] {
      return wf::launch_grad_assemble(h->pts, n, Q, false, F, 6 * q, dinc + order[c0 + q], acc, st);
    }));
  return WF_OK;
}

// ---------------------------------------------------------------------------
// FP64 tree (gradients): the FP32 FFT path's steps with every pattern, table
// and twiddle in FP64 (wf_tree64.hip).  Levels 0..depth all kept; the root
// transition evaluates level 1 on the output grid.
struct Tree64 {
  std::vector<int64_t> toff;  // plain translation tables [level][8][nt][np]
  double2* T = nullptr;
  double2 *patA = nullptr, *patB = nullptr, *G = nullptr;
};

wf_status tree64_prepare(wf_handle* h, hipStream_t st, const GroupCtx& gc, int F, Tree64& tr) {
  const int depth = h->depth;
  tr.toff.assign(depth + 1, 0);
  int64_t tt = 0;
  for (int l = 0; l < depth; ++l) {
    tr.toff[l] = tt;
    tt += 8LL * gc.g[l].nt * gc.g[l].np;
  }
  WF_TRY(dalloc_t(h, "t64_trans", (size_t)std::max<int64_t>(tt, 1), &tr.T));
  for (int l = 0; l < depth; ++l) {
    const Grid gl = gc.g[l];
    const double hc[3] = {0.5 * (h->box[3] - h->box[0]) / std::ldexp(1.0, l + 1),
                          0.5 * (h->box[4] - h->box[1]) / std::ldexp(1.0, l + 1),
                          0.5 * (h->box[5] - h->box[2]) / std::ldexp(1.0, l + 1)};
    const double k0 = gc.k0;
    WF_TRY(timed(h, st, WF_K_TRANS, 16.0 * 8 * gl.nt * gl.np, [&] {
      return wf::launch_trans64(tr.T + tr.toff[l], gl.nt, gl.np, theta_origin(l, gl.nt), hc, k0,
                                st);
    }));
  }
  size_t pat_max = 16, g_max = 16;
  for (int l = 1; l <= depth; ++l)
    pat_max = std::max(pat_max, (size_t)h->levels[l].count * F * gc.g[l].nt * gc.g[l].np);
  for (int l = 0; l < depth; ++l)
    g_max = std::max(g_max, (size_t)h->levels[l + 1].count * F * gc.g[l + 1].nt * gc.g[l].np);
  WF_TRY(dalloc_t(h, "t64_patA", pat_max, &tr.patA));
  WF_TRY(dalloc_t(h, "t64_patB", pat_max, &tr.patB));
  WF_TRY(dalloc_t(h, "t64_G", g_max, &tr.G));
  return WF_OK;
}

wf::Fft64Plan fft64_or_die(wf_handle* h, int n, wf_status* err) {
  wf::Fft64Plan p{};
  if (*err == WF_OK) *err = get_fft64(h, n, &p);
  return p;
}

// FP64 forward of one chunk: currents (FP64) -> leaf sums -> levels -> F0.
wf_status forward_chunk64(wf_handle* h, hipStream_t st, const GroupCtx& gc, const WfIncDev* dinc,
                          const int* order, int c0, int S, int NT, int NP, Tree64& tr,
                          double2* F0) {
  const int depth = h->depth;
  const int64_t n = h->n;
  const int F = 6 * S;
  const float* px = h->pts;
  double2* C = nullptr;
  WF_TRY(dalloc_t(h, "t64_C", (size_t)n * F, &C));
  for (int q = 0; q < S; ++q)
    WF_TRY(timed(h, st, WF_K_CURRENTS, (28.0 + 96.0) * n, [&] {
      return wf::launch_currents(px, px + n, px + 2 * n, px + 3 * n, px + 4 * n, px + 5 * n,
                                 px + 6 * n, n, dinc + order[c0 + q], 1, 1, C + 6 * q, F, 0, st);
    }));
  double2* cur = (depth % 2) ? tr.patA : tr.patB;
  const LevelDev& LL = h->levels[depth];
  wf::Leaf64Args la;
  la.px = px;
  la.py = px + n;
  la.pz = px + 2 * n;
  la.C = C;
  la.F = F;
  la.leaf_begin = LL.begin;
  la.leaf_end = LL.end;
  la.leaf_key = LL.key;
  la.nleaves = LL.count;
  la.depth = depth;
  for (int q = 0; q < 3; ++q) {
    la.box_lo[q] = h->box[q];
    la.leaf_size[q] = (h->box[q + 3] - h->box[q]) / std::ldexp(1.0, depth);
  }
  la.k0 = gc.k0;
  la.nt = gc.g[depth].nt;
  la.np = gc.g[depth].np;
  la.out = cur;
  const double dleaf = (double)la.nt * la.np;
  WF_TRY(timed(h, st, WF_K_LEAF, 8.0 * dleaf * n * F, [&] { return wf::launch_leaf64(la, st); }));
  for (int l = depth - 1; l >= 0; --l) {
    if (cancelled(h)) return set_err(WF_ERR_CANCELLED, "cancelled");
    const Grid gcg = gc.g[l + 1], gp = gc.g[l];
    wf_status err = WF_OK;
    const wf::Fft64Plan pc = fft64_or_die(h, gcg.np, &err), pp = fft64_or_die(h, gp.np, &err),
                        tc = fft64_or_die(h, 2 * gcg.nt, &err),
                        tp = fft64_or_die(h, 2 * gp.nt, &err);
    WF_TRY(err);
    const int nchild = h->levels[l + 1].count, npar = h->levels[l].count;
    const double off_c = theta_origin(l + 1, gcg.nt), off_p = theta_origin(l, gp.nt);
    const int mirror = mirror_of(gp.nt, off_p);
    WF_TRY(timed(h, st, WF_K_PHI, 16.0 * nchild * F * gcg.nt * ((double)gcg.np + gp.np), [&] {
      return wf::launch_phi64(cur, tr.G, (int64_t)nchild * F, gcg.nt, gcg.np, gp.np, pc, pp, st);
    }));
    double2* nxt = (l == 0) ? F0 : ((cur == tr.patA) ? tr.patB : tr.patA);
    const int np_out = (l == 0) ? 1 : npar;
    WF_TRY(timed(h, st, WF_K_THETA, 16.0 * F * ((double)nchild * gcg.nt * gp.np + (double)np_out * gp.nt * gp.np), [&] {
      return wf::launch_theta64(tr.G, F, gcg.nt, gp.np, gp.nt, np_out, h->levels[l].cb,
                                h->levels[l].ce, h->levels[l + 1].key, tr.T + tr.toff[l],
                                off_p - off_c, mirror, tc, tp, nxt, st);
    }));
    cur = nxt;
  }
  return WF_OK;
}

// FP64 reverse of one chunk: gF [F][D] -> leaf cogradients -> Q -> acc.
wf_status backward_chunk64(wf_handle* h, hipStream_t st, const GroupCtx& gc, const WfIncDev* dinc,
                           const int* order, int c0, int S, Tree64& tr, const double2* gF,
                           double* acc) {
  const int depth = h->depth;
  const int64_t n = h->n;
  const int F = 6 * S;
  const double2* gcur = gF;  // cogradient of the current level's patterns
  for (int l = 0; l < depth; ++l) {
    const Grid gp = gc.g[l], gch = gc.g[l + 1];
    wf_status err = WF_OK;
    const wf::Fft64Plan pc = fft64_or_die(h, gch.np, &err), pp = fft64_or_die(h, gp.np, &err),
                        tc = fft64_or_die(h, 2 * gch.nt, &err),
                        tp = fft64_or_die(h, 2 * gp.nt, &err);
    WF_TRY(err);
    const int nchild = h->levels[l + 1].count;
    const double off_c = theta_origin(l + 1, gch.nt), off_p = theta_origin(l, gp.nt);
    const int mirror = mirror_of(gp.nt, off_p);
    WF_TRY(timed(h, st, WF_K_THETA, 0.0, [&] {
      return wf::launch_theta_adj64(gcur, F, gch.nt, gp.np, gp.nt, nchild,
                                    l == 0 ? nullptr : h->levels[l + 1].parent,
                                    h->levels[l + 1].key, tr.T + tr.toff[l], off_p - off_c,
                                    mirror, tp, tc, tr.G, st);
    }));
    double2* nxt = (gcur == tr.patA) ? tr.patB : tr.patA;
    WF_TRY(timed(h, st, WF_K_PHI, 0.0, [&] {
      return wf::launch_phi_adj64(tr.G, nxt, (int64_t)nchild * F, gch.nt, gch.np, gp.np, pp, pc,
                                  st);
    }));
    gcur = nxt;
  }
  double2* Q = nullptr;
  WF_TRY(dalloc_t(h, "t64_Q", (size_t)n * F * 4, &Q));
  const LevelDev& LL = h->levels[depth];
  wf::LeafAdj64Args aa;
  aa.px = h->pts;
  aa.py = h->pts + n;
  aa.pz = h->pts + 2 * n;
  aa.leaf_begin = LL.begin;
  aa.leaf_end = LL.end;
  aa.leaf_key = LL.key;
  aa.nleaves = LL.count;
  aa.depth = depth;
  for (int q = 0; q < 3; ++q) {
    aa.box_lo[q] = h->box[q];
    aa.leaf_size[q] = (h->box[q + 3] - h->box[q]) / std::ldexp(1.0, depth);
  }
  aa.k0 = gc.k0;
  aa.nt = gc.g[depth].nt;
  aa.np = gc.g[depth].np;
  aa.F = F;
  aa.g = gcur;
  aa.Q = Q;
  WF_TRY(timed(h, st, WF_K_LEAF, 0.0, [&] { return wf::launch_leaf_adj64(aa, st); }));
  for (int q = 0; q < S; ++q)
    WF_TRY(timed(h, st, WF_K_CURRENTS, 0.0, [&] {
      return wf::launch_grad_assemble(h->pts, n, Q, true, F, 6 * q, dinc + order[c0 + q], acc, st);
    }));
  return WF_OK;
}

// The launch sequence of one wf_solve (replayable: it allocates only on first
// use and records nothing but kernels while a graph is being captured).
// One chunk of a solve: solves [c0, c0 + S) (in lambda order) of the
// wavelength group [i0, i1), all at the group's depth.
struct Chunk {
  int i0, i1, c0, S;
};

// Runs chunks in order on handle w and stream st; a chunk whose group differs
// from the previous one prepares that group's grids and tables first (in w's
// workspace).
wf_status run_chunks(wf_handle* w, hipStream_t st, const std::vector<Chunk>& cs,
                     const wf_incidence* inc, int NT, int NP, const WfIncDev* dinc,
                     const int* order, float* I_dev, float* E_dev) {
  const int64_t n = w->n;
  const int64_t D = (int64_t)NT * NP;
  const float* px = w->pts;
  const double centre[3] = {0.5 * (w->box[0] + w->box[3]), 0.5 * (w->box[1] + w->box[4]),
                            0.5 * (w->box[2] + w->box[5])};
  const double epi_bytes = (96.0 + 4.0 + (E_dev ? 24.0 : 0.0)) * D;
  GroupCtx gc;
  int cur_group = -1;
  for (const Chunk& ch : cs) {
    if (ch.i0 != cur_group) {
      w->depth = choose_depth(w, inc[order[ch.i0]].lambda_m, NT, NP);
      w->stats.depth = w->depth;
      gc = GroupCtx();
      WF_TRY(prepare_group(w, st, inc[order[ch.i0]].lambda_m, NT, NP, gc));
      cur_group = ch.i0;
    }
    const int depth = w->depth;
    const int c0 = ch.c0, S = ch.S;
    const int F = 6 * S;
    int fps = 6;  // fields per solve of F0: 6, or 2 transverse components
    if (cancelled(w)) return set_err(WF_ERR_CANCELLED, "cancelled");
    double2* F0 = nullptr;
    WF_TRY(dalloc_t(w, "F0", (size_t)D * F, &F0));
    if (depth == 0) {
      double2 *C = nullptr, *part = nullptr;
      WF_TRY(dalloc_t(w, "direct_C", (size_t)n * F, &C));
      const int64_t part_elems = (int64_t)wf::direct_split(D) * F * D;
      WF_TRY(dalloc_t(w, "direct_part", (size_t)part_elems, &part));
      for (int q = 0; q < S; ++q)
        WF_TRY(timed(w, st, WF_K_CURRENTS, (28.0 + 96.0) * n, [&] {
          return wf::launch_currents(px, px + n, px + 2 * n, px + 3 * n, px + 4 * n, px + 5 * n,
                                     px + 6 * n, n, dinc + order[c0 + q], 1, 1, C + 6 * q, F, 0, st);
        }));
      const double k0 = gc.k0;
      WF_TRY(timed(w, st, WF_K_DIRECT, 48.0 * n * D * S, [&] {
        return wf::launch_direct_fp64(px, px + n, px + 2 * n, n, C, F, centre, k0, NT, NP, F0,
                                      part, part_elems, st);
      }));
    } else {
      TreeBufs tb;
      WF_TRY(tree_buffers(w, gc, F, NP, tb));
      WF_TRY(forward_chunk(w, st, gc, dinc, order, c0, S, NT, NP, tb, F0, true, &fps));
    }
    for (int q = 0; q < S; ++q) {
      const int slot = order[c0 + q];
      WF_TRY(timed(w, st, WF_K_EPILOGUE, epi_bytes, [&] {
        if (fps == 2)
          return wf::launch_intensity_transverse(F0 + (int64_t)q * 2 * D, 1, NT, NP,
                                                 I_dev + (int64_t)slot * D,
                                                 E_dev ? E_dev + (int64_t)slot * 6 * D : nullptr,
                                                 nullptr, nullptr, st);
        return wf::launch_intensity(F0 + (int64_t)q * 6 * D, 1, NT, NP, I_dev + (int64_t)slot * D,
                                    E_dev ? E_dev + (int64_t)slot * 6 * D : nullptr, nullptr,
                                    nullptr, st);
      }));
    }
  }
  return WF_OK;
}

// Worker handles for concurrent chunks: created on first use with the
// parent's options (sequential themselves), then bound to the parent's
// geometry and tuning before every solve.
wf_status ensure_workers(wf_handle* h, int K) {
  while ((int)h->subs.size() < K - 1) {
    wf_opts o = h->opts;
    o.use_graph = 0;
    o.solve_streams = 1;
    wf_handle* w = nullptr;
    WF_TRY(wf_create(&w, &o));
    hipStream_t s = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
      wf_destroy(w);
      return set_err(WF_ERR_HIP, "worker stream creation failed");
    }
    h->subs.push_back(w);
    h->sub_st.push_back(s);
  }
  while ((int)h->sub_ev.size() < K) {
    hipEvent_t e = nullptr;
    WF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    h->sub_ev.push_back(e);
  }
  for (int k = 0; k < K - 1; ++k) {
    wf_handle* w = h->subs[k];
    w->n = h->n;
    w->depth = h->depth;
    w->depth_max = h->depth_max;
    w->auto_depth = h->auto_depth;
    w->force_depth = h->force_depth;
    w->tune = h->tune;
    w->tuned_gen = h->tuned_gen;
    w->tuned_lam = h->tuned_lam;
    w->tuned_depth = h->tuned_depth;
    for (int q = 0; q < 6; ++q) {
      w->boxf[q] = h->boxf[q];
      w->box[q] = h->box[q];
    }
    w->pts = h->pts;
    w->perm = h->perm;
    w->keys = h->keys;
    w->levels = h->levels;
    w->geom_gen = h->geom_gen;
    w->has_geom = true;
  }
  return WF_OK;
}

wf_status solve_core(wf_handle* h, hipStream_t st, const wf_incidence* inc, int n_inc, int NT,
                     int NP, const WfIncDev* dinc, const int* order, float* I_dev, float* E_dev) {
  // chunks: per wavelength group, runs of at most max_solves solves
  std::vector<Chunk> chunks;
  for (int i0 = 0; i0 < n_inc;) {
    int i1 = i0 + 1;
    while (i1 < n_inc && inc[order[i1]].lambda_m == inc[order[i0]].lambda_m) ++i1;
    const int depth = choose_depth(h, inc[order[i0]].lambda_m, NT, NP);
    int max_solves = h->opts.field_chunk > 0 ? std::max(1, h->opts.field_chunk / 6)
                     : (h->tune.chunk_solves > 0 ? h->tune.chunk_solves : 8);
    max_solves = std::min(max_solves, 8);
    if (depth == 0) max_solves = std::min(max_solves, 4);
    for (int c0 = i0; c0 < i1; c0 += max_solves)
      chunks.push_back(Chunk{i0, i1, c0, std::min(max_solves, i1 - c0)});
    i0 = i1;
  }
  int K = h->opts.solve_streams > 0 ? h->opts.solve_streams : 4;
  if (h->capturing || h->sequential || h->opts.use_graph) K = 1;  // graphs record one stream
  K = std::min<int>(K, (int)chunks.size());
  if (K <= 1) return run_chunks(h, st, chunks, inc, NT, NP, dinc, order, I_dev, E_dev);
  WF_TRY(ensure_workers(h, K));
  // chunk j goes to worker j mod K (worker 0 is h on the caller's stream);
  // the workers start after everything the caller's stream holds so far
  std::vector<std::vector<Chunk>> part(K);
  for (size_t j = 0; j < chunks.size(); ++j) part[j % K].push_back(chunks[j]);
  WF_HIP(hipEventRecord(h->sub_ev[0], st));
  // the caller's stream waits for every worker that was given work, also when
  // a worker or the handle itself stops early (error, cancellation)
  wf_status err = WF_OK;
  int started = 0;
  for (int k = 1; k < K && err == WF_OK; ++k) {
    prof_reset(h->subs[k - 1]);  // a worker created by this call starts with no records
    if (hipStreamWaitEvent(h->sub_st[k - 1], h->sub_ev[0], 0) != hipSuccess) {
      err = set_err(WF_ERR_HIP, "worker stream wait failed");
      break;
    }
    started = k;
    err = run_chunks(h->subs[k - 1], h->sub_st[k - 1], part[k], inc, NT, NP, dinc, order, I_dev,
                     E_dev);
    if (hipEventRecord(h->sub_ev[k], h->sub_st[k - 1]) != hipSuccess && err == WF_OK)
      err = set_err(WF_ERR_HIP, "worker join event failed");
  }
  if (err == WF_OK) err = run_chunks(h, st, part[0], inc, NT, NP, dinc, order, I_dev, E_dev);
  for (int k = 1; k <= started; ++k) {
    const hipError_t e = hipStreamWaitEvent(st, h->sub_ev[k], 0);
    if (e != hipSuccess && err == WF_OK) err = set_err(WF_ERR_HIP, "join wait failed");
  }
  return err;
}

}  // namespace

extern "C" {

wf_status wf_solve(wf_handle* h, const wf_incidence* inc, int32_t n_inc, wf_grid grid,
                   float* intensity, float* E_cplx, int32_t out_memspace, void* stream) {
  if (!h) return set_err(WF_ERR_INVALID_ARG, "handle is null");
  if (!h->has_geom) return set_err(WF_ERR_NO_GEOMETRY, "wf_solve: no geometry (call wf_set_geometry)");
  if (!intensity) return set_err(WF_ERR_INVALID_ARG, "intensity is null");
  if (grid.n_theta < 1 || grid.n_phi < 2) return set_err(WF_ERR_INVALID_ARG, "grid too small");
  if (out_memspace != WF_MEM_HOST && out_memspace != WF_MEM_DEVICE)
    return set_err(WF_ERR_INVALID_ARG, "out_memspace unknown");
  WF_TRY(check_inc(inc, n_inc));
  WF_TRY(set_device(h));
  hipStream_t st = as_stream(stream);
  const int NT = grid.n_theta, NP = grid.n_phi;
  const int64_t D = (int64_t)NT * NP;
  if (h->depth_max > 0) WF_TRY(check_tree_grid(grid));
  float* I_dev = intensity;
  float* E_dev = E_cplx;
  if (out_memspace == WF_MEM_HOST) {
    WF_TRY(dalloc_t(h, "out_I", (size_t)n_inc * D, &I_dev));
    if (E_cplx) WF_TRY(dalloc_t(h, "out_E", (size_t)n_inc * 6 * D, &E_dev));
  }
  // graph replay: same geometry, incidences, grid and output buffers
  const bool graph_ok = h->opts.use_graph && st != nullptr;
  uint64_t sig = 1469598103934665603ULL;
  auto mix = [&](const void* p, size_t nb) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < nb; ++i) sig = (sig ^ c[i]) * 1099511628211ULL;
  };
  if (graph_ok) {
    mix(&h->geom_gen, sizeof h->geom_gen);
    mix(&n_inc, sizeof n_inc);
    mix(&grid, sizeof grid);
    mix(inc, sizeof(wf_incidence) * n_inc);
    mix(&I_dev, sizeof I_dev);
    mix(&E_dev, sizeof E_dev);
    mix(&h->tune, sizeof h->tune);
    mix(&h->auto_depth, sizeof h->auto_depth);
    if (h->gcache.valid && h->gcache.sig == sig && h->gcache.ws_gen == h->ws_gen) {
      WfIncDev* gdinc = nullptr;
      WF_TRY(upload_incs(h, inc, n_inc, &gdinc, st, "inc_graph"));
      if (h->gcache.ws_gen == h->ws_gen) {  // the upload did not reallocate
        const int d_last = h->stats.depth;
        prof_reset(h);
        h->stats.depth = d_last;
        h->stats_final = true;
        WF_HIP(hipGraphLaunch(h->gcache.exec, st));
        if (out_memspace == WF_MEM_HOST) {
          WF_HIP(hipMemcpyAsync(intensity, I_dev, sizeof(float) * n_inc * D, hipMemcpyDeviceToHost, st));
          if (E_cplx)
            WF_HIP(hipMemcpyAsync(E_cplx, E_dev, sizeof(float) * n_inc * 6 * D, hipMemcpyDeviceToHost, st));
          WF_HIP(hipStreamSynchronize(st));
        }
        return WF_OK;
      }
    }
  }
  WfIncDev* dinc = nullptr;
  WF_TRY(upload_incs(h, inc, n_inc, &dinc, st, graph_ok ? "inc_graph" : "inc"));
  std::vector<int> order(n_inc);
  for (int i = 0; i < n_inc; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return inc[a].lambda_m < inc[b].lambda_m; });
  prof_reset(h);
  WF_TRY(prof_event(h, st, &h->ev_t0));
  WF_TRY(solve_core(h, st, inc, n_inc, NT, NP, dinc, order.data(), I_dev, E_dev));
  WF_TRY(prof_event(h, st, &h->ev_t1));
  if (graph_ok) {
    // capture the same launch sequence (every buffer now exists) for replays
    h->gcache.reset();
    const uint64_t ws_before = h->ws_gen;
    WF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    h->capturing = true;
    const wf_status cs = solve_core(h, st, inc, n_inc, NT, NP, dinc, order.data(), I_dev, E_dev);
    h->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(st, &g);
    if (cs != WF_OK) {
      if (g) hipGraphDestroy(g);
      return cs;
    }
    if (ce != hipSuccess) return set_err(WF_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
    if (h->ws_gen != ws_before) {
      // a buffer moved while recording: earlier nodes may hold its old
      // address, so the graph is dropped (the eager results stand)
      hipGraphDestroy(g);
      g = nullptr;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = g ? hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) : hipSuccess;
    if (ie != hipSuccess) {
      hipGraphDestroy(g);
      return set_err(WF_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(ie));
    }
    if (g) {
      h->gcache.graph = g;
      h->gcache.exec = ex;
      h->gcache.sig = sig;
      h->gcache.ws_gen = h->ws_gen;
      h->gcache.valid = true;
    }
  }
  if (out_memspace == WF_MEM_HOST) {
    WF_HIP(hipMemcpyAsync(intensity, I_dev, sizeof(float) * n_inc * D, hipMemcpyDeviceToHost, st));
    if (E_cplx)
      WF_HIP(hipMemcpyAsync(E_cplx, E_dev, sizeof(float) * n_inc * 6 * D, hipMemcpyDeviceToHost, st));
    WF_HIP(hipStreamSynchronize(st));
  }
  return WF_OK;
}

wf_status wf_po_vjp(wf_handle* h, const wf_incidence* inc, int32_t n_inc, wf_grid grid,
                    const float* grad_I, float* intensity, float* grad_x, float* grad_n,
                    float* grad_a, int32_t fp64, int32_t memspace, void* stream) {
  if (!h) return set_err(WF_ERR_INVALID_ARG, "handle is null");
  if (!h->has_geom) return set_err(WF_ERR_NO_GEOMETRY, "wf_po_vjp: no geometry");
  // grad_I null: the forward intensity only, in the arithmetic of the reverse pass
  const bool fwd_only = grad_I == nullptr;
  if (fwd_only && !intensity) return set_err(WF_ERR_INVALID_ARG, "grad_I and intensity are null");
  if (grid.n_theta < 1 || grid.n_phi < 2) return set_err(WF_ERR_INVALID_ARG, "grid too small");
  if (memspace != WF_MEM_HOST && memspace != WF_MEM_DEVICE)
    return set_err(WF_ERR_INVALID_ARG, "memspace unknown");
  WF_TRY(check_inc(inc, n_inc));
  WF_TRY(set_device(h));
  hipStream_t st = as_stream(stream);
  const int64_t n = h->n;
  if (fp64 < 0 || fp64 > 2) return set_err(WF_ERR_INVALID_ARG, "wf_po_vjp: fp64 must be 0, 1 or 2");
  const bool direct = fp64 == 1 || h->depth_max == 0;
  const bool tree64 = !direct && fp64 == 2;
  const int NT = grid.n_theta, NP = grid.n_phi;
  const int64_t D = (int64_t)NT * NP;
  if (!direct) WF_TRY(check_tree_grid(grid));
  WfIncDev* dinc = nullptr;
  WF_TRY(upload_incs(h, inc, n_inc, &dinc, st));
  const float* gI = grad_I;
  float* I_dev = intensity;
  float *gx = grad_x, *gn = grad_n, *ga = grad_a;
  if (memspace == WF_MEM_HOST) {
    if (!fwd_only) {
      float* t = nullptr;
      WF_TRY(dalloc_t(h, "vjp_gI", (size_t)n_inc * D, &t));
      WF_HIP(hipMemcpyAsync(t, grad_I, sizeof(float) * n_inc * D, hipMemcpyHostToDevice, st));
      gI = t;
    }
    if (intensity) WF_TRY(dalloc_t(h, "vjp_I", (size_t)n_inc * D, &I_dev));
    if (grad_x) WF_TRY(dalloc_t(h, "vjp_gx", (size_t)3 * n, &gx));
    if (grad_n) WF_TRY(dalloc_t(h, "vjp_gn", (size_t)3 * n, &gn));
    if (grad_a) WF_TRY(dalloc_t(h, "vjp_ga", (size_t)n, &ga));
  }
  double* acc = nullptr;
  WF_TRY(dalloc_t(h, "vjp_acc", (size_t)7 * n, &acc));
  WF_HIP(hipMemsetAsync(acc, 0, sizeof(double) * 7 * n, st));
  std::vector<int> order(n_inc);
  for (int i = 0; i < n_inc; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return inc[a].lambda_m < inc[b].lambda_m; });
  const double centre[3] = {0.5 * (h->box[0] + h->box[3]), 0.5 * (h->box[1] + h->box[4]),
                            0.5 * (h->box[2] + h->box[5])};
  // The reverse kernels (theta_adj, phi_adj) are the exact adjoints of the
  // FP32 LDS FFT level transitions, so the recomputed forward of a tree
  // gradient runs those transitions (tune.agg_kernel = 1) whatever the
  // handle's setting, and its leaf sum on FP32 MFMA: the gradient is then the
  // adjoint of the forward it pairs with, up to FP32 rounding.
  struct TuneGuard {
    wf_handle* h;
    wf_tune t;
    int32_t prec;
    ~TuneGuard() {
      h->tune = t;
      h->opts.precision = prec;
    }
  } guard{h, h->tune, h->opts.precision};
  h->tune.beta = 0.0;  // gradients sample at opts.beta, whatever the forward was tuned to
  if (!direct) {
    h->tune.agg_kernel = 1;
    if (h->opts.precision == WF_PREC_BF16X3) h->opts.precision = WF_PREC_FP32;
  }
  prof_reset(h);
  WF_TRY(prof_event(h, st, &h->ev_t0));
  const float* px = h->pts;
  const int max_solves = direct ? 1 : 8;
  for (int i0 = 0; i0 < n_inc;) {
    int i1 = i0 + 1;
    while (i1 < n_inc && inc[order[i1]].lambda_m == inc[order[i0]].lambda_m) ++i1;
    GroupCtx gc;
    h->depth = direct ? 0 : choose_depth(h, inc[order[i0]].lambda_m, NT, NP);
    h->stats.depth = h->depth;
    if (direct) {
      gc.lam = inc[order[i0]].lambda_m;
      gc.k0 = 2.0 * WF_PI / gc.lam;
    } else {
      WF_TRY(prepare_group(h, st, inc[order[i0]].lambda_m, NT, NP, gc, false));
    }
    for (int c0 = i0; c0 < i1; c0 += max_solves) {
      const int S = std::min(max_solves, i1 - c0);
      const int F = 6 * S;
      if (cancelled(h)) return set_err(WF_ERR_CANCELLED, "cancelled");
      double2 *F0 = nullptr, *gF = nullptr;
      WF_TRY(dalloc_t(h, "F0", (size_t)D * F, &F0));
      WF_TRY(dalloc_t(h, "vjp_gF", (size_t)D * F, &gF));
      TreeBufs tb;
      Tree64 t64;
      if (tree64) {
        WF_TRY(tree64_prepare(h, st, gc, F, t64));
        WF_TRY(forward_chunk64(h, st, gc, dinc, order.data(), c0, S, NT, NP, t64, F0));
      } else if (direct) {
        double2 *C = nullptr, *part = nullptr;
        WF_TRY(dalloc_t(h, "direct_C", (size_t)n * F, &C));
        const int64_t part_elems = (int64_t)wf::direct_split(D) * F * D;
        WF_TRY(dalloc_t(h, "direct_part", (size_t)part_elems, &part));
        WF_HIP(wf::launch_currents(px, px + n, px + 2 * n, px + 3 * n, px + 4 * n, px + 5 * n,
                                   px + 6 * n, n, dinc + order[c0], 1, 1, C, F, 0, st));
        const double k0 = gc.k0;
        WF_TRY(timed(h, st, WF_K_DIRECT, 48.0 * n * D * S, [&] {
          return wf::launch_direct_fp64(px, px + n, px + 2 * n, n, C, F, centre, k0, NT, NP, F0,
                                        part, part_elems, st);
        }));
      } else {
        WF_TRY(tree_buffers(h, gc, F, NP, tb));
        int fps6 = 6;  // gradients: the six fields (the reverse pass needs them)
        WF_TRY(forward_chunk(h, st, gc, dinc, order.data(), c0, S, NT, NP, tb, F0, false, &fps6));
      }
      if (fwd_only) {
        for (int q = 0; q < S; ++q) {
          const int slot = order[c0 + q];
          WF_TRY(timed(h, st, WF_K_EPILOGUE, 0.0, [&] {
            return wf::launch_intensity(F0 + (int64_t)q * 6 * D, 1, NT, NP,
                                        I_dev + (int64_t)slot * D, nullptr, nullptr, nullptr, st);
          }));
        }
        continue;
      }
      for (int q = 0; q < S; ++q) {
        const int slot = order[c0 + q];
        WF_TRY(timed(h, st, WF_K_EPILOGUE, 0.0, [&] {
          return wf::launch_epilogue_adj(F0 + (int64_t)q * 6 * D, gI + (int64_t)slot * D, NT, NP,
                                         gF + (int64_t)q * 6 * D,
                                         I_dev ? I_dev + (int64_t)slot * D : nullptr, st);
        }));
      }
      if (direct) {
        double2* Qd = nullptr;
        WF_TRY(dalloc_t(h, "adj_Qd", (size_t)n * 24, &Qd));
        const double k0 = gc.k0;
        WF_TRY(timed(h, st, WF_K_DIRECT, 0.0, [&] {
          return wf::launch_direct_adj_fp64(px, n, centre, k0, NT, NP, gF, Qd, st);
        }));
        WF_TRY(timed(h, st, WF_K_CURRENTS, 0.0, [&] {
          return wf::launch_grad_assemble(px, n, Qd, true, 6, 0, dinc + order[c0], acc, st);
        }));
      } else if (tree64) {
        WF_TRY(backward_chunk64(h, st, gc, dinc, order.data(), c0, S, t64, gF, acc));
      } else {
        WF_TRY(backward_chunk(h, st, gc, dinc, order.data(), c0, S, NT, NP, tb, gF, acc));
      }
    }
    i0 = i1;
  }
  if (!fwd_only) WF_HIP(wf::launch_grad_scatter(acc, h->perm, n, gx, gn, ga, st));
  WF_TRY(prof_event(h, st, &h->ev_t1));
  if (memspace == WF_MEM_HOST) {
    if (intensity)
      WF_HIP(hipMemcpyAsync(intensity, I_dev, sizeof(float) * n_inc * D, hipMemcpyDeviceToHost, st));
    if (grad_x) WF_HIP(hipMemcpyAsync(grad_x, gx, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, st));
    if (grad_n) WF_HIP(hipMemcpyAsync(grad_n, gn, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, st));
    if (grad_a) WF_HIP(hipMemcpyAsync(grad_a, ga, sizeof(float) * n, hipMemcpyDeviceToHost, st));
    WF_HIP(hipStreamSynchronize(st));
  }
  return WF_OK;
}

wf_status wf_po_solve(wf_handle* h, const wf_surface* s, const wf_incidence* inc, int32_t n_inc,
                      wf_grid grid, int32_t depth, float* intensity, float* E_cplx, void* stream) {
  if (!s) return set_err(WF_ERR_INVALID_ARG, "surface is null");
  WF_TRY(wf_set_geometry(h, s, depth, stream));
  return wf_solve(h, inc, n_inc, grid, intensity, E_cplx, s->memspace, stream);
}

wf_status wf_get_stats(wf_handle* h, wf_stats* out) {
  if (!h || !out) return set_err(WF_ERR_INVALID_ARG, "null argument");
  WF_TRY(set_device(h));
  WF_TRY(prof_finalize(h));
  *out = h->stats;
  return WF_OK;
}

wf_status wf_geometry_info(wf_handle* h, int64_t* n, int32_t* depth_built, double* box6) {
  if (!h) return set_err(WF_ERR_INVALID_ARG, "handle is null");
  if (!h->has_geom) return set_err(WF_ERR_NO_GEOMETRY, "no geometry set");
  if (n) *n = h->n;
  if (depth_built) *depth_built = h->depth_max;
  if (box6)
    for (int q = 0; q < 6; ++q) box6[q] = h->box[q];
  return WF_OK;
}

wf_status wf_get_tune(wf_handle* h, wf_tune* out) {
  if (!h || !out) return set_err(WF_ERR_INVALID_ARG, "null argument");
  *out = h->tune;
  return WF_OK;
}

wf_status wf_set_tune(wf_handle* h, const wf_tune* t) {
  if (!h || !t) return set_err(WF_ERR_INVALID_ARG, "null argument");
  if (!(std::isfinite(t->leaf_s_per_pair) && t->leaf_s_per_pair >= 0 &&
        std::isfinite(t->agg_s_per_byte) && t->agg_s_per_byte >= 0 && std::isfinite(t->level_s) &&
        t->level_s >= 0))
    return set_err(WF_ERR_INVALID_ARG, "wf_tune: coefficients must be finite and >= 0");
  if (t->chunk_solves < 0 || t->chunk_solves > 8)
    return set_err(WF_ERR_INVALID_ARG, "wf_tune: chunk_solves must be in [0, 8]");
  if (t->agg_kernel < 0 || t->agg_kernel > 1 || t->agg_fft_rows < 0 || t->agg_fused < 0 ||
      t->agg_fused > 2 || t->theta_tile < 0 || t->theta_tile > 3 ||
      (t->phi_rows != 0 && t->phi_rows != 16 && t->phi_rows != 32) ||
      !(t->beta >= 0 && t->beta < 100))
    return set_err(WF_ERR_INVALID_ARG,
                   "wf_tune: agg_kernel must be 0 or 1, agg_fused 0, 1 or 2, agg_fft_rows >= 0, "
                   "theta_tile 0..3, phi_rows 0, 16 or 32, 0 <= beta < 100");
  h->tune = *t;
  h->tune.leaf_quad = (t->leaf_quad >= 0 && t->leaf_quad <= 3) ? t->leaf_quad : 1;
  return WF_OK;
}

wf_status wf_autotune(wf_handle* h, const wf_incidence* inc, int32_t n_inc, wf_grid grid,
                      int32_t* best_depth, void* stream) {
  if (!h) return set_err(WF_ERR_INVALID_ARG, "handle is null");
  if (!h->has_geom) return set_err(WF_ERR_NO_GEOMETRY, "wf_autotune: no geometry");
  WF_TRY(check_inc(inc, n_inc));
  WF_TRY(set_device(h));
  if (h->depth_max == 0) {
    if (best_depth) *best_depth = 0;
    return WF_OK;
  }
  WF_TRY(check_tree_grid(grid));
  hipStream_t st = as_stream(stream);
  const int NT = grid.n_theta, NP = grid.n_phi;
  const int64_t D = (int64_t)NT * NP;
  // the first wavelength group, at most 8 solves
  std::vector<wf_incidence> sub;
  for (int i = 0; i < n_inc && sub.size() < 8; ++i)
    if (inc[i].lambda_m == inc[0].lambda_m) sub.push_back(inc[i]);
  const int S = (int)sub.size();
  std::vector<int> order(S);
  for (int i = 0; i < S; ++i) order[i] = i;
  WfIncDev* dinc = nullptr;
  WF_TRY(upload_incs(h, sub.data(), S, &dinc, st, "inc_tune"));
  float* I_dev = nullptr;
  WF_TRY(dalloc_t(h, "tune_I", (size_t)S * D, &I_dev));
  const double lam_s = sample_lambda(h, sub[0].lambda_m);
  const wf_tune saved = h->tune;
  struct Run {
    int d;
    double total, leaf, agg, other;
  };
  // eager solve at a forced depth: one untimed pass (allocations), one timed
  auto run = [&](int d, Run& r) -> wf_status {
    h->force_depth = d;
    h->sequential = true;  // per-kernel times without other streams' launches beside them
    for (int rep = 0; rep < 2; ++rep) {
      prof_reset(h);
      wf_status e = prof_event(h, st, &h->ev_t0);
      if (e == WF_OK) e = solve_core(h, st, sub.data(), S, NT, NP, dinc, order.data(), I_dev, nullptr);
      if (e == WF_OK) e = prof_event(h, st, &h->ev_t1);
      if (e == WF_OK) e = prof_finalize(h);
      if (e != WF_OK) {
        h->force_depth = -1;
        h->sequential = false;
        return e;
      }
    }
    h->force_depth = -1;
    h->sequential = false;
    const double* k = h->stats.kern_ms;
    r.d = d;
    r.total = 1e-3 * h->stats.ms_total;
    r.leaf = 1e-3 * k[WF_K_LEAF];
    r.agg = 1e-3 * (k[WF_K_PHI] + k[WF_K_THETA] + k[WF_K_TRANS]);
    r.other = 1e-3 * (k[WF_K_CURRENTS] + k[WF_K_EPILOGUE]);
    return WF_OK;
  };
  // candidate depths: the model's three cheapest
  std::vector<std::pair<double, int>> rank;
  auto rank_depths = [&]() {
    rank.clear();
    for (int d = 1; d <= h->depth_max; ++d)
      rank.push_back({model_cost(h, h->tune, d, lam_s, NT, NP), d});
    std::sort(rank.begin(), rank.end());
    if (rank.size() > 3) rank.resize(3);
  };
  rank_depths();
  // sampling constant: the smallest of 3 and 2 below opts.beta whose
  // intensities stay within kBetaGate of the opts.beta solve, every solve
  wf_status e = WF_OK;
  if (h->opts.ref_dtheta_m <= 0 && h->opts.beta > 2.0) {
    constexpr double kBetaGate = 1e-4;
    std::vector<float> Iref((size_t)S * D), Ib((size_t)S * D);
    auto fetch = [&](std::vector<float>& dst) -> wf_status {
      WF_HIP(hipMemcpyAsync(dst.data(), I_dev, sizeof(float) * dst.size(), hipMemcpyDeviceToHost,
                            st));
      WF_HIP(hipStreamSynchronize(st));
      return WF_OK;
    };
    Run r;
    h->tune.beta = 0.0;
    e = run(rank[0].second, r);
    if (e == WF_OK) e = fetch(Iref);
    double chosen = 0.0;
    for (const double b : {3.0, 2.0}) {
      if (e != WF_OK || b >= h->opts.beta) continue;
      h->tune.beta = b;
      e = run(rank[0].second, r);
      if (e == WF_OK) e = fetch(Ib);
      if (e != WF_OK) break;
      double worst = 0.0;
      for (int i = 0; i < S; ++i) {
        double num = 0.0, den = 0.0;
        for (int64_t q = 0; q < D; ++q) {
          const double a = Ib[(size_t)i * D + q], c = Iref[(size_t)i * D + q];
          num += (a - c) * (a - c);
          den += c * c;
        }
        worst = std::max(worst, den > 0 ? std::sqrt(num / den) : 0.0);
      }
      if (!(worst <= kBetaGate)) break;
      chosen = b;
    }
    if (e != WF_OK) {
      h->tune = saved;
      return e;
    }
    h->tune.beta = chosen;
    rank_depths();
  }
  // leaf kernel and solves per tree pass at the model's best depth (a BF16x3
  // choice): quad-symmetric kernel with 4 or 8 solves per pass, or the plain
  // kernel with 8
  if (h->opts.precision == WF_PREC_BF16X3 && S > 4 && h->opts.field_chunk <= 0) {
    static const int32_t cand[3][2] = {{1, 4}, {1, 8}, {0, 8}};  // leaf_quad, chunk_solves
    double tbest = 0;
    int ibest = 0;
    for (int c = 0; c < 3; ++c) {
      h->tune.leaf_quad = cand[c][0];
      h->tune.chunk_solves = cand[c][1];
      Run r;
      e = run(rank[0].second, r);
      if (e != WF_OK) {
        h->tune = saved;
        return e;
      }
      if (c == 0 || r.total < tbest) {
        tbest = r.total;
        ibest = c;
      }
    }
    h->tune.leaf_quad = cand[ibest][0];
    h->tune.chunk_solves = cand[ibest][1];
  }
  // level-transition kernels at the model's best depth: the staged fused
  // form or two passes and the two-pass theta tile of long rows (dense MFMA
  // levels), then how many of the largest levels run the LDS FFT kernels;
  // every candidate gives the same result to FP32 rounding
  if (h->tune.agg_kernel == 0) {
    static const int32_t cand[3][2] = {{2, 0}, {0, 0}, {2, 2}};  // agg_fused, theta_tile
    double tbest = 0;
    int ibest = 0;
    for (int c = 0; c < 3; ++c) {
      h->tune.agg_fused = cand[c][0];
      h->tune.theta_tile = cand[c][1];
      Run r;
      e = run(rank[0].second, r);
      if (e != WF_OK) {
        h->tune = saved;
        return e;
      }
      if (c == 0 || r.total < tbest) {
        tbest = r.total;
        ibest = c;
      }
    }
    h->tune.agg_fused = cand[ibest][0];
    h->tune.theta_tile = cand[ibest][1];
    // the extended-row threshold above which transitions take the LDS FFT
    // kernels instead of the MFMA ones: from the current setting down, each
    // candidate moving the next smaller level grid to the FFT; kept while
    // measured faster
    std::vector<int> rows2;  // extended rows of the parent grids of the transitions
    for (int l = 0; l < rank[0].second; ++l) {
      const int r2 = 2 * ((l == 0) ? NT : level_grid(h, l, lam_s).nt);
      if (r2 <= h->tune.agg_fft_rows) rows2.push_back(r2);
    }
    std::sort(rows2.rbegin(), rows2.rend());
    rows2.erase(std::unique(rows2.begin(), rows2.end()), rows2.end());
    int32_t best_rows = h->tune.agg_fft_rows;
    for (size_t i = 0; i < rows2.size() && i < 3; ++i) {
      h->tune.agg_fft_rows = rows2[i] - 1;
      Run r;
      e = run(rank[0].second, r);
      if (e != WF_OK) {
        h->tune = saved;
        return e;
      }
      if (!(r.total < tbest)) break;
      tbest = r.total;
      best_rows = rows2[i] - 1;
    }
    h->tune.agg_fft_rows = best_rows;
  }
  std::vector<Run> runs;
  for (const auto& c : rank) {
    Run r;
    e = run(c.second, r);
    if (e != WF_OK) {
      h->tune = saved;
      return e;
    }
    runs.push_back(r);
  }
  // fit: leaf seconds per pair, aggregation seconds per byte, per-level rest
  double sl = 0, pl = 0, sa = 0, ba = 0, rest = 0;
  int nrest = 0;
  const Run* best = &runs[0];
  for (const Run& r : runs) {
    const Grid gd = level_grid(h, r.d, lam_s);
    sl += r.leaf;
    pl += (double)h->n * gd.nt * gd.np * S;
    wf_tune unit = h->tune;
    unit.leaf_s_per_pair = 0.0;
    unit.agg_s_per_byte = 1.0;
    unit.level_s = 0.0;
    sa += r.agg;
    ba += model_cost(h, unit, r.d, lam_s, NT, NP) * S;
    rest += std::max(0.0, r.total - r.leaf - r.agg - r.other) / ((double)r.d * S);
    ++nrest;
    if (r.total < best->total) best = &r;
  }
  if (pl > 0) h->tune.leaf_s_per_pair = sl / pl;
  if (ba > 0) h->tune.agg_s_per_byte = sa / ba;
  if (nrest) h->tune.level_s = rest / nrest;
  h->tune.calibrated = 1;
  h->tuned_gen = h->geom_gen;
  h->tuned_lam = lam_s;
  h->tuned_depth = best->d;
  if (best_depth) *best_depth = best->d;
  return WF_OK;
}

}  // extern "C"

// ============================================================ batches =====
namespace {

// A batch segment's geometry: the sorted points and the per-level node lists
// of its octree, compacted into one device allocation.
struct SegGeom {
  int64_t n = 0;
  int depth_max = 0;
  bool auto_depth = false;
  float boxf[6];
  double box[6];
  float* pts = nullptr;
  int32_t* perm = nullptr;
  uint32_t* keys = nullptr;
  std::vector<LevelDev> levels;
  void* mem = nullptr;
  uint64_t gen = 0;
  std::vector<GraphCache> graphs;  // one per worker
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int last_depth = 0;
  bool timed = false;
};

uint64_t g_seg_gen = 0;

}  // namespace

struct wf_batch {
  wf_opts opts;
  int device = 0;
  std::vector<wf_handle*> workers;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> done;
  std::vector<SegGeom> segs;
  bool tuned = false;
  void* out = nullptr;  // device staging of host outputs
  size_t out_cap = 0;
};

namespace {

// Points the handle at a segment's geometry (no copies).
void bind_segment(wf_handle* h, const SegGeom& g) {
  h->n = g.n;
  h->depth = g.depth_max;
  h->depth_max = g.depth_max;
  h->auto_depth = g.auto_depth;
  for (int q = 0; q < 6; ++q) {
    h->boxf[q] = g.boxf[q];
    h->box[q] = g.box[q];
  }
  h->pts = g.pts;
  h->perm = g.perm;
  h->keys = g.keys;
  h->levels = g.levels;
  h->geom_gen = g.gen;
  h->has_geom = true;
}

void free_segment(SegGeom& g) {
  for (auto& c : g.graphs) c.reset();
  if (g.ev0) hipEventDestroy(g.ev0);
  if (g.ev1) hipEventDestroy(g.ev1);
  if (g.mem) hipFree(g.mem);
  g.mem = nullptr;
  g.ev0 = g.ev1 = nullptr;
}

}  // namespace

extern "C" {

wf_status wf_batch_create(wf_batch** out, const wf_opts* opts, int32_t n_streams) {
  if (!out) return set_err(WF_ERR_INVALID_ARG, "wf_batch_create: out is null");
  *out = nullptr;
  if (n_streams < 1 || n_streams > 16)
    return set_err(WF_ERR_INVALID_ARG, "wf_batch_create: n_streams must be in [1, 16]");
  wf_batch* b = new wf_batch();
  for (int w = 0; w < n_streams; ++w) {
    wf_handle* h = nullptr;
    const wf_status e = wf_create(&h, opts);
    if (e != WF_OK) {
      wf_batch_destroy(b);
      return e;
    }
    h->opts.solve_streams = 1;  // the batch's own streams give the concurrency
    b->workers.push_back(h);
    if (w == 0) {
      b->opts = h->opts;
      b->device = h->device;
    }
    hipStream_t s = nullptr;
    hipEvent_t ev = nullptr;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
      if (s) hipStreamDestroy(s);
      wf_batch_destroy(b);
      return set_err(WF_ERR_HIP, "wf_batch_create: stream/event creation failed");
    }
    b->streams.push_back(s);
    b->done.push_back(ev);
  }
  *out = b;
  return WF_OK;
}

void wf_batch_destroy(wf_batch* b) {
  if (!b) return;
  hipSetDevice(b->device);
  for (auto s : b->streams) hipStreamSynchronize(s);
  for (auto& g : b->segs) free_segment(g);
  if (b->out) hipFree(b->out);
  for (auto ev : b->done) hipEventDestroy(ev);
  for (auto s : b->streams) hipStreamDestroy(s);
  for (auto h : b->workers) wf_destroy(h);
  delete b;
}

wf_status wf_batch_add(wf_batch* b, const wf_surface* s, int32_t depth, int32_t* index) {
  if (!b) return set_err(WF_ERR_INVALID_ARG, "batch is null");
  wf_handle* h = b->workers[0];
  hipStream_t st = b->streams[0];
  WF_TRY(wf_set_geometry(h, s, depth, st));
  SegGeom g;
  g.n = h->n;
  g.depth_max = h->depth_max;
  g.auto_depth = h->auto_depth;
  for (int q = 0; q < 6; ++q) {
    g.boxf[q] = h->boxf[q];
    g.box[q] = h->box[q];
  }
  const int64_t n = h->n;
  size_t nodes = 0;
  for (const LevelDev& L : h->levels) nodes += (size_t)L.count;
  const size_t bytes = sizeof(float) * 7 * n + 2 * sizeof(int32_t) * n + 6 * sizeof(int32_t) * nodes +
                       2 * sizeof(int32_t) * h->levels.size();
  WF_HIP(hipMalloc(&g.mem, bytes));
  char* p = static_cast<char*>(g.mem);
  auto take = [&](size_t nb) {
    char* q = p;
    p += nb;
    return q;
  };
  g.pts = reinterpret_cast<float*>(take(sizeof(float) * 7 * n));
  g.perm = reinterpret_cast<int32_t*>(take(sizeof(int32_t) * n));
  g.keys = reinterpret_cast<uint32_t*>(take(sizeof(uint32_t) * n));
  wf_status e = WF_OK;
  auto cp = [&](void* dst, const void* src, size_t nb) {
    if (e == WF_OK && nb && hipMemcpyAsync(dst, src, nb, hipMemcpyDeviceToDevice, st) != hipSuccess)
      e = set_err(WF_ERR_HIP, "wf_batch_add: device copy failed");
  };
  cp(g.pts, h->pts, sizeof(float) * 7 * n);
  cp(g.perm, h->perm, sizeof(int32_t) * n);
  cp(g.keys, h->keys, sizeof(uint32_t) * n);
  g.levels.resize(h->levels.size());
  for (size_t l = 0; l < h->levels.size(); ++l) {
    const LevelDev& L = h->levels[l];
    LevelDev& G = g.levels[l];
    const size_t c = (size_t)L.count;
    G.count = L.count;
    G.key = reinterpret_cast<uint32_t*>(take(4 * c));
    G.begin = reinterpret_cast<int32_t*>(take(4 * c));
    G.end = reinterpret_cast<int32_t*>(take(4 * c));
    G.cb = reinterpret_cast<int32_t*>(take(4 * c));
    G.ce = reinterpret_cast<int32_t*>(take(4 * c));
    G.parent = reinterpret_cast<int32_t*>(take(4 * c));
    G.span = reinterpret_cast<int32_t*>(take(8));
    cp(G.span, L.span, 8);
    cp(G.key, L.key, 4 * c);
    cp(G.begin, L.begin, 4 * c);
    cp(G.end, L.end, 4 * c);
    cp(G.cb, L.cb, 4 * c);
    cp(G.ce, L.ce, 4 * c);
    cp(G.parent, L.parent, 4 * c);
  }
  if (e == WF_OK && hipStreamSynchronize(st) != hipSuccess)
    e = set_err(WF_ERR_HIP, "wf_batch_add: stream synchronize failed");
  h->has_geom = false;
  if (e != WF_OK) {
    hipFree(g.mem);
    return e;
  }
  g.gen = (1ULL << 63) | ++g_seg_gen;
  g.graphs.resize(b->workers.size());
  if (hipEventCreate(&g.ev0) != hipSuccess || hipEventCreate(&g.ev1) != hipSuccess) {
    free_segment(g);
    return set_err(WF_ERR_HIP, "wf_batch_add: event creation failed");
  }
  if (index) *index = (int32_t)b->segs.size();
  b->segs.push_back(std::move(g));
  return WF_OK;
}

// Queues every segment's solve on the worker streams (the body of
// wf_batch_solve; the caller joins the workers whatever this returns).
static wf_status batch_queue(wf_batch* b, const wf_incidence* inc, int32_t n_inc, wf_grid grid, float* out,
                      int64_t per_seg) {
  const int nseg = (int)b->segs.size();
  const int nw = (int)b->workers.size();
  // largest segments first; each goes to the worker with the least queued
  // points (deterministic, so a segment keeps its worker and its graph)
  std::vector<int> order(nseg);
  for (int i = 0; i < nseg; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int c) { return b->segs[a].n > b->segs[c].n; });
  // cost-model calibration on the largest auto-depth segment
  if (!b->tuned) {
    for (int i : order) {
      if (!b->segs[i].auto_depth) continue;
      wf_handle* h0 = b->workers[0];
      bind_segment(h0, b->segs[i]);
      std::swap(h0->gcache, b->segs[i].graphs[0]);
      const wf_status e = wf_autotune(h0, inc, n_inc, grid, nullptr, b->streams[0]);
      std::swap(h0->gcache, b->segs[i].graphs[0]);
      h0->has_geom = false;
      WF_TRY(e);
      wf_tune t;
      WF_TRY(wf_get_tune(h0, &t));
      for (int w = 1; w < nw; ++w) WF_TRY(wf_set_tune(b->workers[w], &t));
      b->tuned = true;
      break;
    }
  }
  std::vector<double> load(nw, 0.0);
  double lam_max = inc[0].lambda_m;
  for (int i = 1; i < n_inc; ++i) lam_max = std::max(lam_max, inc[i].lambda_m);
  for (int i : order) {
    SegGeom& g = b->segs[i];
    const int w = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[w] += (double)g.n;
    wf_handle* h = b->workers[w];
    hipStream_t st = b->streams[w];
    bind_segment(h, g);
    std::swap(h->gcache, g.graphs[w]);
    WF_HIP(hipEventRecord(g.ev0, st));
    const wf_status e =
        wf_solve(h, inc, n_inc, grid, out + (int64_t)i * per_seg, nullptr, WF_MEM_DEVICE, st);
    std::swap(h->gcache, g.graphs[w]);
    g.last_depth = choose_depth(h, lam_max, grid.n_theta, grid.n_phi);
    h->has_geom = false;
    WF_TRY(e);
    WF_HIP(hipEventRecord(g.ev1, st));
    g.timed = true;
  }
  return WF_OK;
}

wf_status wf_batch_solve(wf_batch* b, const wf_incidence* inc, int32_t n_inc, wf_grid grid,
                         float* intensity, int32_t out_memspace, void* stream) {
  if (!b) return set_err(WF_ERR_INVALID_ARG, "batch is null");
  if (!intensity) return set_err(WF_ERR_INVALID_ARG, "intensity is null");
  if (b->segs.empty()) return set_err(WF_ERR_NO_GEOMETRY, "wf_batch_solve: no segments");
  if (out_memspace != WF_MEM_HOST && out_memspace != WF_MEM_DEVICE)
    return set_err(WF_ERR_INVALID_ARG, "out_memspace unknown");
  if (grid.n_theta < 1 || grid.n_phi < 2) return set_err(WF_ERR_INVALID_ARG, "grid too small");
  WF_TRY(check_inc(inc, n_inc));
  WF_HIP(hipSetDevice(b->device));
  const int nseg = (int)b->segs.size();
  const int nw = (int)b->workers.size();
  const int64_t per_seg = (int64_t)n_inc * grid.n_theta * grid.n_phi;
  for (const SegGeom& g : b->segs)
    if (g.depth_max > 0) WF_TRY(check_tree_grid(grid));
  float* out = intensity;
  if (out_memspace == WF_MEM_HOST) {
    const size_t need = sizeof(float) * (size_t)per_seg * nseg;
    if (b->out_cap < need) {
      if (b->out) {
        for (auto s : b->streams) WF_HIP(hipStreamSynchronize(s));
        hipFree(b->out);
      }
      b->out = nullptr;
      b->out_cap = 0;
      WF_HIP(hipMalloc(&b->out, need));
      b->out_cap = need;
    }
    out = static_cast<float*>(b->out);
  }
  hipStream_t caller = as_stream(stream);
  // the workers start after the caller's prior work (device outputs)
  WF_HIP(hipEventRecord(b->done[0], caller));
  for (int w = 0; w < nw; ++w) WF_HIP(hipStreamWaitEvent(b->streams[w], b->done[0], 0));
  // from here on every exit joins all worker streams back into the caller's
  // stream: segments queued before an error or a cancellation keep writing
  // into `out`, so the caller's stream must not run ahead of them
  wf_status err = batch_queue(b, inc, n_inc, grid, out, per_seg);
  for (int w = 0; w < nw; ++w) {
    const hipError_t e1 = hipEventRecord(b->done[w], b->streams[w]);
    const hipError_t e2 = (e1 == hipSuccess) ? hipStreamWaitEvent(caller, b->done[w], 0) : e1;
    if (e2 != hipSuccess && err == WF_OK)
      err = set_err(WF_ERR_HIP, "batch join failed: %s", hipGetErrorString(e2));
  }
  if (err != WF_OK && out_memspace == WF_MEM_HOST) hipStreamSynchronize(caller);
  WF_TRY(err);

  if (out_memspace == WF_MEM_HOST) {
    WF_HIP(hipMemcpyAsync(intensity, out, sizeof(float) * (size_t)per_seg * nseg,
                          hipMemcpyDeviceToHost, caller));
    WF_HIP(hipStreamSynchronize(caller));
  }
  return WF_OK;
}

wf_status wf_heightfield_points(const float* height, int32_t numel, int32_t numelvert, float period,
                                float unitz, int32_t start, int32_t end, float* out_soa,
                                int32_t memspace, void* stream) {
  if (!height || !out_soa) return set_err(WF_ERR_INVALID_ARG, "height or out_soa is null");
  if (numel < 2 || numelvert < 1)
    return set_err(WF_ERR_INVALID_ARG, "heightfield needs numel >= 2 and numelvert >= 1");
  if (start < 0 || end < start || end >= numel)
    return set_err(WF_ERR_INVALID_ARG, "arc window [%d, %d] outside [0, %d)", start, end, numel);
  if ((int64_t)numel * numelvert >= (1LL << 31) - 1)
    return set_err(WF_ERR_INVALID_ARG, "heightfield too large");
  if (!(std::isfinite(period) && std::isfinite(unitz)))
    return set_err(WF_ERR_INVALID_ARG, "period and unitz must be finite");
  if (memspace != WF_MEM_HOST && memspace != WF_MEM_DEVICE)
    return set_err(WF_ERR_INVALID_ARG, "memspace unknown");
  hipStream_t st = as_stream(stream);
  const int64_t nh = 2LL * numel * numelvert;
  const int64_t n = (int64_t)(end - start + 1) * numelvert;
  // temporaries are stream-ordered (no device-wide synchronisation)
  float* dh = const_cast<float*>(height);
  float* dout = out_soa;
  int* bad = nullptr;
  hipError_t e = hipSuccess;
  if (memspace == WF_MEM_HOST) {
    dh = nullptr;
    dout = nullptr;
    e = hipMallocAsync(reinterpret_cast<void**>(&dh), sizeof(float) * nh, st);
    if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&dout), sizeof(float) * 7 * n, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(dh, height, sizeof(float) * nh, hipMemcpyHostToDevice, st);
  }
  if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&bad), sizeof(int), st);
  int hbad = 0x7fffffff;
  if (e == hipSuccess)
    e = wf::launch_heightfield_points(dh, numel, numelvert, period, unitz, start, end, dout, bad,
                                      st);
  if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && memspace == WF_MEM_HOST)
    e = hipMemcpyAsync(out_soa, dout, sizeof(float) * 7 * n, hipMemcpyDeviceToHost, st);
  if (bad) (void)hipFreeAsync(bad, st);
  if (memspace == WF_MEM_HOST) {
    if (dh) (void)hipFreeAsync(dh, st);
    if (dout) (void)hipFreeAsync(dout, st);
  }
  const hipError_t es = hipStreamSynchronize(st);
  if (e == hipSuccess) e = es;
  if (e != hipSuccess) return set_err(WF_ERR_HIP, "wf_heightfield_points: %s", hipGetErrorString(e));
  if (hbad != 0x7fffffff)
    return set_err(WF_ERR_INVALID_ARG, "heightfield vertex at x = y = 0 (sample j = %d, i = %d)",
                   hbad % numel, hbad / numel);
  return WF_OK;
}

// Argument checks of the 2D cross-section solve and the per-incidence device
// parameters (incidence basis, k, far-field factors) in hinc[n_inc].
static wf_status rough2d_params(const double* x, const double* y, int32_t numel,
                                const wf_rough2d_incidence* inc, int32_t n_inc, int32_t n_phi,
                                double distance_m, double radius_m, const double* intensity,
                                int32_t memspace, wf::Rough2dIncDev* hinc) {
  if (!x || !y || !inc || !intensity) return set_err(WF_ERR_INVALID_ARG, "null argument");
  if (numel < 2 || numel > (1 << 26)) return set_err(WF_ERR_INVALID_ARG, "numel must be in [2, 2^26]");
  if (n_inc < 1 || n_inc > 65535) return set_err(WF_ERR_INVALID_ARG, "n_inc must be in [1, 65535]");
  if (n_phi < 1 || n_phi > (1 << 24)) return set_err(WF_ERR_INVALID_ARG, "n_phi must be in [1, 2^24]");
  if ((int64_t)numel * n_inc > (1LL << 28))
    return set_err(WF_ERR_INVALID_ARG, "numel * n_inc too large");
  if (!(distance_m > 0 && std::isfinite(distance_m)) || !(radius_m > 0 && std::isfinite(radius_m)))
    return set_err(WF_ERR_INVALID_ARG, "distance and radius must be positive and finite");
  if (memspace != WF_MEM_HOST && memspace != WF_MEM_DEVICE)
    return set_err(WF_ERR_INVALID_ARG, "memspace unknown");
  const double mu0 = 4e-7 * WF_PI, eps0 = 8.8541878e-12, c0 = 299792458.0;  // rough2d.cu:25-28
  for (int i = 0; i < n_inc; ++i) {
    const wf_rough2d_incidence& a = inc[i];
    if (!(a.lambda_m > 0 && std::isfinite(a.lambda_m)))
      return set_err(WF_ERR_INVALID_ARG, "incidence %d: wavelength must be positive", i);
    if (!(std::isfinite(a.theta) && std::fabs(a.theta) < WF_PI / 2 && std::isfinite(a.phi)))
      return set_err(WF_ERR_INVALID_ARG, "incidence %d: need |theta| < pi/2 and finite phi", i);
    if (a.pol != 'M' && a.pol != 'E')
      return set_err(WF_ERR_INVALID_ARG, "incidence %d: pol must be 'M' or 'E'", i);
    if (!(std::isfinite(a.eta_re) && std::isfinite(a.eta_im)))
      return set_err(WF_ERR_INVALID_ARG, "incidence %d: eta must be finite", i);
    wf::Rough2dIncDev& d = hinc[(size_t)i];
    const double ct = std::cos(a.theta), st_ = std::sin(a.theta);
    const double cp = std::cos(a.phi), sp = std::sin(a.phi);
    d.w[0] = ct * cp;  // rough2d.cu:264-266
    d.w[1] = ct * sp;
    d.w[2] = st_;
    const double v[3] = {-st_ * cp, -st_ * sp, ct};
    double* A = (a.pol == 'M') ? d.Ed : d.Hd;  // rough2d.cu:268-277
    double* B = (a.pol == 'M') ? d.Hd : d.Ed;
    for (int c = 0; c < 3; ++c) A[c] = v[c];
    if (a.pol == 'M') {  // H = w x E (:292)
      B[0] = d.w[1] * v[2] - d.w[2] * v[1];
      B[1] = d.w[2] * v[0] - d.w[0] * v[2];
      B[2] = d.w[0] * v[1] - d.w[1] * v[0];
    } else {  // E = H x w (:296)
      B[0] = v[1] * d.w[2] - v[2] * d.w[1];
      B[1] = v[2] * d.w[0] - v[0] * d.w[2];
      B[2] = v[0] * d.w[1] - v[1] * d.w[0];
    }
    d.k = 2.0 * WF_PI / a.lambda_m;
    d.krho = d.k * ct;  // :474
    d.ct = ct;
    d.st = st_;
    d.z0 = std::sqrt(mu0 / eps0);
    d.eta_re = a.eta_re;
    d.eta_im = a.eta_im;
    const double omega = d.k * c0;
    d.fe = -omega * mu0 / 4.0;  // :82-90
    d.fh = -omega * eps0 / 4.0;
  }
  return WF_OK;
}

wf_status wf_rough2d_solve(const double* x, const double* y, int32_t numel,
                           const wf_rough2d_incidence* inc, int32_t n_inc, int32_t n_phi,
                           double distance_m, double radius_m, double* intensity,
                           int32_t memspace, void* stream) {
  if (n_inc < 1 || n_inc > 65535) return set_err(WF_ERR_INVALID_ARG, "n_inc must be in [1, 65535]");
  std::vector<wf::Rough2dIncDev> hinc((size_t)n_inc);
  WF_TRY(rough2d_params(x, y, numel, inc, n_inc, n_phi, distance_m, radius_m, intensity, memspace,
                        hinc.data()));
  hipStream_t st = as_stream(stream);
  const int64_t nI = (int64_t)n_inc * n_phi;
  double* dx = const_cast<double*>(x);
  double* dy = const_cast<double*>(y);
  double* dI = intensity;
  wf::Rough2dIncDev* dinc = nullptr;
  double2* C = nullptr;
  int* bad = nullptr;
  hipError_t e = hipSuccess;
  if (memspace == WF_MEM_HOST) {
    dx = dy = dI = nullptr;
    e = hipMallocAsync(reinterpret_cast<void**>(&dx), sizeof(double) * numel, st);
    if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&dy), sizeof(double) * numel, st);
    if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&dI), sizeof(double) * nI, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(double) * numel, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dy, y, sizeof(double) * numel, hipMemcpyHostToDevice, st);
  }
  if (e == hipSuccess)
    e = hipMallocAsync(reinterpret_cast<void**>(&dinc), sizeof(wf::Rough2dIncDev) * n_inc, st);
  if (e == hipSuccess)
    e = hipMallocAsync(reinterpret_cast<void**>(&C), sizeof(double2) * 6 * (size_t)numel * n_inc, st);
  if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&bad), sizeof(int), st);
  const int P = wf::rough2d_split(numel, n_inc, n_phi);
  double2* part = nullptr;
  if (e == hipSuccess)
    e = hipMallocAsync(reinterpret_cast<void**>(&part), sizeof(double2) * 6 * (size_t)P * n_inc * n_phi,
                       st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dinc, hinc.data(), sizeof(wf::Rough2dIncDev) * n_inc,
                       hipMemcpyHostToDevice, st);
  int hbad = 0x7fffffff;
  if (e == hipSuccess)
    e = wf::launch_rough2d(dx, dy, numel, dinc, n_inc, n_phi, distance_m, radius_m, C, part, P, dI,
                           bad, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess && memspace == WF_MEM_HOST)
    e = hipMemcpyAsync(intensity, dI, sizeof(double) * nI, hipMemcpyDeviceToHost, st);
  if (bad) (void)hipFreeAsync(bad, st);
  if (part) (void)hipFreeAsync(part, st);
  if (C) (void)hipFreeAsync(C, st);
  if (dinc) (void)hipFreeAsync(dinc, st);
  if (memspace == WF_MEM_HOST) {
    if (dx) (void)hipFreeAsync(dx, st);
    if (dy) (void)hipFreeAsync(dy, st);
    if (dI) (void)hipFreeAsync(dI, st);
  }
  const hipError_t es = hipStreamSynchronize(st);
  if (e == hipSuccess) e = es;
  if (e != hipSuccess) return set_err(WF_ERR_HIP, "wf_rough2d_solve: %s", hipGetErrorString(e));
  if (hbad != 0x7fffffff)
    return set_err(WF_ERR_INVALID_ARG, "cross-section segment %d has zero length or a non-finite vertex",
                   hbad);
  return WF_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free 2D cross-section solve.  Its scratch memory (incidence
// parameters, currents, partial sums, the bad-segment flag) lives in one
// file-scope device array; every call on a device shares that device's copy,
// so calls that overlapped would race: a lock held until the call's stream has
// finished serialises them.  In host code the array's name
// stands for a host-side symbol, so kernels receive the device address that
// hipGetSymbolAddress returns, looked up once per device.
static __device__ __align__(256) unsigned char wf_r2_scratch[WF_ROUGH2D_NOALLOC_SCRATCH_BYTES];
static void* g_r2_scratch_ptr[64] = {nullptr};
static std::mutex g_r2_mutex;  // the host staging buffer and the per-device lookup
static wf::Rough2dIncDev g_r2_hinc[WF_ROUGH2D_NOALLOC_MAX_INC];

wf_status wf_rough2d_solve_noalloc(const double* x, const double* y, int32_t numel,
                                   const wf_rough2d_incidence* inc, int32_t n_inc, int32_t n_phi,
                                   double distance_m, double radius_m, double* intensity,
                                   void* stream) {
  if (n_inc < 1 || n_inc > WF_ROUGH2D_NOALLOC_MAX_INC)
    return set_err(WF_ERR_INVALID_ARG, "wf_rough2d_solve_noalloc: n_inc must be in [1, %d]",
                   WF_ROUGH2D_NOALLOC_MAX_INC);
  std::lock_guard<std::mutex> lock(g_r2_mutex);
  WF_TRY(rough2d_params(x, y, numel, inc, n_inc, n_phi, distance_m, radius_m, intensity,
                        WF_MEM_DEVICE, g_r2_hinc));
  // scratch layout: [incidences | currents | partial sums | flag], 256 B aligned
  auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const int P = wf::rough2d_split(numel, n_inc, n_phi);
  const size_t b_inc = up(sizeof(wf::Rough2dIncDev) * (size_t)n_inc);
  const size_t b_cur = up(sizeof(double2) * 6 * (size_t)numel * n_inc);
  const size_t b_part = up(sizeof(double2) * 6 * (size_t)P * n_inc * n_phi);
  const size_t need = b_inc + b_cur + b_part + 256;
  if (need > (size_t)WF_ROUGH2D_NOALLOC_SCRATCH_BYTES)
    return set_err(WF_ERR_INVALID_ARG,
                   "wf_rough2d_solve_noalloc: needs %zu bytes of scratch (%d samples x %d "
                   "incidences x %d azimuths), the maximum is %d",
                   need, numel, n_inc, n_phi, WF_ROUGH2D_NOALLOC_SCRATCH_BYTES);
  int dev = 0;
  WF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return set_err(WF_ERR_INVALID_ARG, "device ordinal %d unsupported", dev);
  if (!g_r2_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(wf_r2_scratch));
    if (e != hipSuccess || !ptr)
      return set_err(WF_ERR_HIP, "wf_rough2d_solve_noalloc: scratch lookup failed: %s",
                     hipGetErrorString(e));
    g_r2_scratch_ptr[dev] = ptr;
  }
  unsigned char* base = static_cast<unsigned char*>(g_r2_scratch_ptr[dev]);
  wf::Rough2dIncDev* dinc = reinterpret_cast<wf::Rough2dIncDev*>(base);
  double2* C = reinterpret_cast<double2*>(base + b_inc);
  double2* part = reinterpret_cast<double2*>(base + b_inc + b_cur);
  int* bad = reinterpret_cast<int*>(base + b_inc + b_cur + b_part);
  hipStream_t st = as_stream(stream);
  int hbad = 0x7fffffff;
  hipError_t e = hipMemcpyAsync(dinc, g_r2_hinc, sizeof(wf::Rough2dIncDev) * n_inc,
                                hipMemcpyHostToDevice, st);
  // launch_rough2d clears the flag with a kernel before use
  if (e == hipSuccess)
    e = wf::launch_rough2d(x, y, numel, dinc, n_inc, n_phi, distance_m, radius_m, C, part, P,
                           intensity, bad, st);
  if (e == hipSuccess) e = hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, st);
  const hipError_t es = hipStreamSynchronize(st);
  if (e == hipSuccess) e = es;
  if (e != hipSuccess) return set_err(WF_ERR_HIP, "wf_rough2d_solve_noalloc: %s", hipGetErrorString(e));
  if (hbad != 0x7fffffff)
    return set_err(WF_ERR_INVALID_ARG, "cross-section segment %d has zero length or a non-finite vertex",
                   hbad);
  return WF_OK;
}

wf_status wf_average_grid(const float* in, int32_t n_grids, int32_t n_theta, int32_t n_phi,
                          int32_t n_theta_out, int32_t n_phi_out, float* out, int32_t memspace,
                          void* stream) {
  if (!in || !out) return set_err(WF_ERR_INVALID_ARG, "in or out is null");
  if (n_grids < 1 || n_theta_out < 1 || n_phi_out < 1 || n_theta < n_theta_out ||
      n_phi < n_phi_out)
    return set_err(WF_ERR_INVALID_ARG, "average grid %dx%d -> %dx%d: output must not be finer",
                   n_theta, n_phi, n_theta_out, n_phi_out);
  if (memspace != WF_MEM_HOST && memspace != WF_MEM_DEVICE)
    return set_err(WF_ERR_INVALID_ARG, "memspace unknown");
  hipStream_t st = as_stream(stream);
  const int64_t nin = (int64_t)n_grids * n_theta * n_phi;
  const int64_t nout = (int64_t)n_grids * n_theta_out * n_phi_out;
  const float* din = in;
  float* dout = out;
  hipError_t e = hipSuccess;
  float* tin = nullptr;
  float* tout = nullptr;
  if (memspace == WF_MEM_HOST) {
    e = hipMallocAsync(reinterpret_cast<void**>(&tin), sizeof(float) * nin, st);
    if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&tout), sizeof(float) * nout, st);
    if (e == hipSuccess) e = hipMemcpyAsync(tin, in, sizeof(float) * nin, hipMemcpyHostToDevice, st);
    din = tin;
    dout = tout;
  }
  if (e == hipSuccess)
    e = wf::launch_average_grid(din, n_grids, n_theta, n_phi, n_theta_out, n_phi_out, dout, st);
  if (memspace == WF_MEM_HOST) {
    if (e == hipSuccess)
      e = hipMemcpyAsync(out, tout, sizeof(float) * nout, hipMemcpyDeviceToHost, st);
    if (tin) (void)hipFreeAsync(tin, st);
    if (tout) (void)hipFreeAsync(tout, st);
    const hipError_t es = hipStreamSynchronize(st);
    if (e == hipSuccess) e = es;
  }
  if (e != hipSuccess) return set_err(WF_ERR_HIP, "wf_average_grid: %s", hipGetErrorString(e));
  return WF_OK;
}

wf_status wf_read_heightfield(const char* path, int32_t numel, int32_t numelvert, float* out) {
  if (!path || !out) return set_err(WF_ERR_INVALID_ARG, "path or out is null");
  if (numel < 1 || numelvert < 1) return set_err(WF_ERR_INVALID_ARG, "numel, numelvert >= 1");
  FILE* f = fopen(path, "rb");
  if (!f) return set_err(WF_ERR_INVALID_ARG, "cannot open %s", path);
  const size_t want = 2 * (size_t)numel * (size_t)numelvert;
  const size_t got = fread(out, sizeof(float), want, f);
  fclose(f);
  if (got != want)
    return set_err(WF_ERR_INVALID_ARG, "%s holds %zu of %zu float32 values", path, got, want);
  return WF_OK;
}

wf_status wf_batch_info(wf_batch* b, int32_t* n_segments, int32_t* depth, double* ms,
                        int64_t* n_points) {
  if (!b) return set_err(WF_ERR_INVALID_ARG, "batch is null");
  WF_HIP(hipSetDevice(b->device));
  if (n_segments) *n_segments = (int32_t)b->segs.size();
  for (size_t i = 0; i < b->segs.size(); ++i) {
    const SegGeom& g = b->segs[i];
    if (depth) depth[i] = g.last_depth;
    if (n_points) n_points[i] = g.n;
    if (ms) {
      float t = 0;
      if (g.timed) {
        WF_HIP(hipEventSynchronize(g.ev1));
        WF_HIP(hipEventElapsedTime(&t, g.ev0, g.ev1));
      }
      ms[i] = t;
    }
  }
  return WF_OK;
}

}  // extern "C"
