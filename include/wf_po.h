/*
 * wf_po.h -- C-ABI of the MI355X-native physical-optics (PO) far-field solver.
 *
 * Drop-in boundary for WaveFiber3d's PO path.  The reference has no library
 * API: its "entry point" is the driver main() (fibersim/treePO.cu:21-992,
 * compareMie/sphere.cu:21-966, compareBEM3d/ellipsoid.cu:20-990) calling the
 * header-level functions of util/*.h over caller-owned thrust vectors.  Each
 * declaration below names the reference code it replaces.
 *
 * Conventions
 *  - Plain C types only; every function returns wf_status (never exits; the
 *    reference's checkCudaErrors() -> exit() at util/helper_cuda.h:582-595 is
 *    replaced by status codes plus a thread-local message, wf_last_error()).
 *  - Units are SI (metres), as in the reference (util/util.h:16-19).
 *  - Directions use the reference's latitude convention (treePO.cu:378-391):
 *    s(theta, phi) = (cos t cos p, cos t sin p, sin t); output grids are
 *    theta_j = -pi/2 + 1e-10 + pi j/Nt, phi_k = 2 pi k/Np, flat index j*Np + k.
 *  - Output intensity is float32 [n_inc][Nt][Np] (treePO.cu:953-957 layout).
 *    The complex far field E'' = E_t - Z0 s x H_t (util/fields.h:189-200),
 *    which the reference never stores, is returned as float32 pairs
 *    [n_inc][3][Nt][Np][2]; its global phase uses the exact FP64
 *    e^{-i k0 D}, D = 1 m (treePO.cu:508-509 evaluates it in FP32).
 *  - memspace WF_MEM_DEVICE: inputs and outputs are device pointers and the
 *    call is asynchronous on `stream` (a hipStream_t; NULL = default stream).
 *    WF_MEM_HOST: inputs/outputs are host pointers and the call returns after
 *    the results are in host memory.
 *  - A handle owns all device workspaces and the cached octree of the current
 *    geometry.  One handle per host thread; handles are not shared concurrently.
 */
#ifndef WF_PO_H
#define WF_PO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WF_ABI_VERSION 5

/* depth argument of wf_set_geometry: build the octree to the deepest useful
 * depth (about log8 n) and let each wavelength group of a solve use the
 * depth that the cost model (wf_tune) rates fastest. */
#define WF_DEPTH_AUTO (-1)

typedef enum wf_status {
  WF_OK = 0,
  WF_ERR_INVALID_ARG = 1,     /* bad pointer/size/value; message names the field */
  WF_ERR_GRID_TOO_COARSE = 2, /* strict_grid and output grid coarser than level 1 (reference Q7) */
  WF_ERR_OOM = 3,             /* device allocation failed or workspace cap exceeded */
  WF_ERR_HIP = 4,             /* HIP runtime error (message carries hipGetErrorString) */
  WF_ERR_CANCELLED = 5,       /* *cancel_flag became nonzero between stages */
  WF_ERR_UNSUPPORTED = 6,     /* e.g. an FFT length with a prime factor above 13 */
  WF_ERR_NO_GEOMETRY = 7      /* wf_solve before wf_set_geometry */
} wf_status;

enum { WF_MEM_HOST = 0, WF_MEM_DEVICE = 1 };

/* Leaf-GEMM arithmetic.  All modes accumulate in FP32; phases are range
 * reduced exactly in revolutions.
 *  WF_PREC_FP32       FP32 MFMA (v_mfma_f32_16x16x4_f32)
 *  WF_PREC_BF16X3     split hi/lo BF16 MFMA, 3 products (v_mfma_f32_16x16x32_bf16)
 *  WF_PREC_FP32_VALU  FP32 on the vector ALU (check path for the MFMA kernels) */
enum { WF_PREC_FP32 = 0, WF_PREC_BF16X3 = 1, WF_PREC_FP32_VALU = 2 };

/* Surface samples, SoA float32, n each (compareMie/sphere.cu:92-125 input
 * form; util/tree.h:67-97 Points and :38-62 Area). */
typedef struct wf_surface {
  const float *x, *y, *z;    /* positions (m) */
  const float *nx, *ny, *nz; /* outward unit normals */
  const float* area;         /* area per sample (m^2) */
  int64_t n;
  int32_t memspace; /* WF_MEM_HOST or WF_MEM_DEVICE */
} wf_surface;

/* One solve = one (wavelength, incidence direction, polarization).
 * treePO.cu:501-509 (lambda, k0), :559-601 (theta_i, phi_i, eta, pol, basis). */
typedef struct wf_incidence {
  double theta, phi;    /* incidence latitude / azimuth (rad) */
  double lambda_m;      /* vacuum wavelength (m) */
  char pol;             /* 'M' = TM, 'E' = TE (treePO.cu:573-601) */
  double eta_re, eta_im;/* complex IOR of the body, outside medium = 1 (treePO.cu:566-568) */
  double beam_sigma_m;  /* Gaussian taper exp(-z^2/(2 sigma^2)) (tangentplane.h:33-34);
                           <= 0 or non-finite: taper off */
} wf_incidence;

typedef struct wf_grid {
  int32_t n_theta, n_phi; /* output far-field grid (treePO.cu:63-77) */
} wf_grid;

typedef struct wf_opts {
  int32_t struct_size;   /* sizeof(wf_opts) (ABI guard); wf_default_opts sets it */
  int32_t device;        /* HIP device ordinal; -1 = current device */
  int32_t precision;     /* WF_PREC_* (leaf radiation GEMM) */
  int32_t strict_grid;   /* 1: refuse output grids coarser than level 1, like the
                            reference's upsample-only interpolation (interpolation.h:27-31) */
  double beta;           /* excess-bandwidth constant of the level sampling rule
                            (treePO.cu:92 uses 6 on its own extent definition) */
  double box_enlarge;    /* automatic root box = bbox centre +- enlarge * half extent
                            (treePO.cu:95-98 uses 1.1) */
  double lambda_sample_m;/* wavelength that sets the level grids; 0 = the group's own
                            (the reference fixes 0.4 um, treePO.cu:347) */
  int32_t has_box;       /* 1: use box[] as the root box (treePO.cu:220-230) */
  double box[6];         /* {min x, min y, min z, max x, max y, max z} (m) */
  int64_t workspace_cap_bytes;   /* 0 = unlimited; exceeded -> WF_ERR_OOM */
  volatile int32_t* cancel_flag; /* nullable; polled between stages */
  int32_t field_chunk;   /* max complex fields (6 per solve) per tree pass; 0 = auto */
  int32_t use_graph;     /* 1: capture wf_solve into a hipGraph and replay it when the same
                            geometry, incidences, grid and outputs come back (needs a
                            non-null stream; the first call runs eagerly, then captures) */
  int32_t solve_streams; /* wf_solve spreads its chunks of solves (wavelength groups, and
                            the field chunks of one group) over this many streams, each
                            with its own workspace, forked from and joined back into the
                            caller's stream; results are identical to 1 (sequential).
                            0 = default (4); graph capture and wf_autotune run sequentially.
                            Each extra stream keeps a workspace of its own (as large as the
                            handle's; workspace_cap_bytes applies to each) */
  double ref_dtheta_m;   /* > 0 together with ref_dphi_m: the reference's level sampling
                            rule instead of this library's (treePO.cu:342-351): level i
                            takes N = ceil(k d + beta (k d)^(1/3)), k = 2 pi / lambda_s,
                            with d = 0.5^i ref_dtheta_m for theta and 0.5^i ref_dphi_m
                            for phi (the reference's extent x rho: fiber period x 1.2 and
                            2 max(r1, r2) x 1.2, sphere diameter x 1.3, ellipsoid box
                            extents x 3), rounded up to 5-smooth lengths (phi even).  Use
                            with beta = 6 and lambda_sample_m = 0.4e-6 for the reference's
                            grids; its theta count covers the z extent only, so nodes wider
                            than tall are under-sampled as in the reference */
  double ref_dphi_m;
} wf_opts;

typedef struct wf_handle wf_handle;

/* Kernel kinds of the per-kernel timing table in wf_stats. */
enum {
  WF_K_CURRENTS = 0, /* work unit: algorithmic HBM bytes */
  WF_K_LEAF = 1,     /* work unit: real FLOP (8 per complex MAC) */
  WF_K_PHI = 2,      /* bytes */
  WF_K_THETA = 3,    /* bytes */
  WF_K_TRANS = 4,    /* bytes */
  WF_K_EPILOGUE = 5, /* bytes */
  WF_K_DIRECT = 6,   /* real FLOP (48 per point-direction-solve) */
  WF_K_COUNT = 8
};

/* Timing/roofline breakdown of the last solve, from HIP events recorded around
 * every kernel launch on the solve's stream (milliseconds).  For an
 * asynchronous (device-memspace) solve, wf_get_stats waits for those events. */
typedef struct wf_stats {
  double ms_total;
  double kern_ms[WF_K_COUNT];       /* summed launch durations per kernel kind */
  int32_t kern_launches[WF_K_COUNT];
  double kern_work[WF_K_COUNT];     /* summed algorithmic work (see kinds) */
  int64_t leaf_pairs;               /* sum over lambda groups of D_leaf * N */
  int32_t depth, n_levels;
  int32_t level_nt[16], level_np[16]; /* level grids of the last lambda group */
  int32_t level_nodes[16];            /* active nodes per level */
  /* Level transitions (aggregation) of the last solve: their number and their
   * compulsory HBM bytes -- every child pattern read once and every parent
   * pattern written once (complex64; 16 B per output direction and field
   * for the FP64 root accumulators) -- whatever intermediate buffers and
   * tables the kernels use. */
  int32_t agg_transitions;
  double agg_bytes;
} wf_stats;

int32_t wf_abi_version(void);
const char* wf_last_error(void);
void wf_default_opts(wf_opts* opts);

wf_status wf_create(wf_handle** out, const wf_opts* opts);
void wf_destroy(wf_handle* h);

/* Geometry upload + octree build (replaces treePO.cu:137-337: H2D copies,
 * build_octree_kernel<256> at util/tree.h:289-715, the active lists and
 * per-point centres).  The points are validated (finite, unit-ish normals,
 * non-negative area), Morton-keyed with the reference's float split planes,
 * stably sorted and cut into per-level active-node lists.  depth in [0, 10]
 * or WF_DEPTH_AUTO. */
wf_status wf_set_geometry(wf_handle* h, const wf_surface* s, int32_t depth, void* stream);

/* PO far-field solve on the current geometry (the treePO.cu:501-961 loop body:
 * currents tangentplane.h:84-92, field_times_area fields.h:123,
 * multiply_current_factor fields.h:147, leaf sums fields.h:41-50 +
 * treePO.cu:652-683, copyandinterp interpolation.h:82-92, translation_compute
 * fields.h:92-103, sumchildren fields.h:153-170, computeintensity
 * fields.h:172-204).  intensity [n_inc][Nt][Np] float32; E_cplx nullable,
 * [n_inc][3][Nt][Np][2] float32.  Output memspace as given. */
wf_status wf_solve(wf_handle* h, const wf_incidence* inc, int32_t n_inc, wf_grid grid,
                   float* intensity, float* E_cplx, int32_t out_memspace, void* stream);

/* One-shot: wf_set_geometry + wf_solve (outputs in the surface's memspace). */
wf_status wf_po_solve(wf_handle* h, const wf_surface* s, const wf_incidence* inc, int32_t n_inc,
                      wf_grid grid, int32_t depth, float* intensity, float* E_cplx, void* stream);

/* Plain method in FP64 (depth-0 semantics of sphere.cu:592-658 and
 * :909-923; FP64 precedent compareBEM3d/ellipsoid_simple.cu:229-261): the
 * oracle-on-GPU and the speedup denominator.  Centre = root box centre.
 * intensity [n_inc][Nt][Np] double; E_cplx nullable [n_inc][3][Nt][Np][2] double. */
wf_status wf_po_solve_fp64_direct(wf_handle* h, const wf_surface* s, const wf_incidence* inc,
                                  int32_t n_inc, wf_grid grid, double* intensity, double* E_cplx,
                                  void* stream);

/* Stage entry points (parity tests per stage; all device pointers, async). */

/* computeLocalTangentPlane + field_times_area + multiply_current_factor
 * (tangentplane.h:84-92, fields.h:123-126, :147-150): currents of the surface
 * (in the given order) for n_inc incidences, complex [n_inc][n][6] as float
 * pairs (fp64 = 0) or double pairs (fp64 = 1). */
wf_status wf_currents(wf_handle* h, const wf_surface* s, const wf_incidence* inc, int32_t n_inc,
                      int32_t fp64, void* out, void* stream);

/* build_octree_kernel (tree.h:289-715) output in the reference's form: perm[n]
 * (original index at sorted position), node_begin/node_end for all
 * (8^(depth+1)-1)/7 nodes in level order.  Uses the current geometry. */
wf_status wf_octree_nodes(wf_handle* h, int32_t* perm, int32_t* node_begin, int32_t* node_end,
                          void* stream);

/* copyandinterp/interpFFT (interpolation.h:11-92) for the tree's patterns:
 * `count` child patterns [count][ntc][npc] complex64 on the half-sample theta
 * grid theta_j = -pi/2 + pi (j + 1/2)/ntc are evaluated on the parent grid
 * [ntp][npp] with theta origin off_tp (phi origin 0), using the 2 pi
 * great-circle extension in theta (npp even).  Device pointers. */
wf_status wf_interp(wf_handle* h, int32_t count, int32_t ntc, int32_t npc, const float* child,
                    int32_t ntp, int32_t npp, double off_tp, float* parent, void* stream);

/* Reverse mode (no reference counterpart; for callers that fit geometry to
 * measured patterns).  Vector-Jacobian product of the solve on the current
 * geometry: given grad_I = dL/dI [n_inc][Nt][Np] (float32), returns
 * grad_x = dL/dx_n and grad_n = dL/dn_n ([3][n] each) and grad_a = dL/da_n
 * ([n]), all in the caller's original point order (any may be null), and the
 * forward intensity (nullable).  The forward pass is recomputed; the reverse
 * pass needs no stored per-level fields.  fp64 selects the arithmetic:
 *   1  the FP64 direct sum forward and backward (the reference gradient; also
 *      used on a depth-0 geometry);
 *   2  the tree with every pattern, table and accumulator in FP64 and its
 *      exact adjoint (gradients of log-intensity losses, which weight the
 *      dimmest directions by 1/I, need FP64 patterns to reach ~1e-4);
 *   0  the tree with FP32 patterns: FP32 FFT level transitions and FP32 MFMA
 *      leaf sums (whatever the handle's tune and precision) and their exact
 *      adjoints.
 * grad_I = NULL with intensity non-null computes only the forward intensity
 * in the selected arithmetic, so that a loss and its gradient can be
 * evaluated on the same forward.  Normal derivatives of the Fresnel currents
 * are analytic (chain rule through the tangent-plane formulas); they are
 * one-sided at the lit/shadow boundary and at exact normal incidence. */
wf_status wf_po_vjp(wf_handle* h, const wf_incidence* inc, int32_t n_inc, wf_grid grid,
                    const float* grad_I, float* intensity, float* grad_x, float* grad_n,
                    float* grad_a, int32_t fp64, int32_t memspace, void* stream);

/* Current geometry: point count, built octree depth (depth_max of a
 * WF_DEPTH_AUTO build) and root box {min x, y, z, max x, y, z}; any may be null. */
wf_status wf_geometry_info(wf_handle* h, int64_t* n, int32_t* depth_built, double* box6);

/* Statistics of the last wf_solve on this handle. */
wf_status wf_get_stats(wf_handle* h, wf_stats* out);

/* ---------------------------------------------------------------------------
 * Depth and kernel autotuning (no reference counterpart: the reference takes
 * the depth from its command line, treePO.cu:27-42, and has one leaf kernel).
 *
 * A WF_DEPTH_AUTO geometry is built once to depth_max; a solve of one
 * wavelength group on depth d < depth_max uses levels 0..d of the same tree
 * (the level lists and keys of a depth-d build are prefixes of the deeper
 * one).  The depth minimises
 *   leaf_s_per_pair * n * Nt_d * Np_d
 *   + agg_s_per_byte * sum_{l<d} 48 (nodes_{l+1} Nt_{l+1} Np_{l+1} + nodes_l Nt_l Np_l)
 *   + level_s * d
 * per solve.  wf_autotune fits the three coefficients from timed solves of
 * the current geometry at the model's best candidate depths, picks the fastest
 * of (quad-symmetric BF16x3 leaf kernel, 4 or 8 solves per pass) and (plain
 * kernel, 8 solves per pass), and stores the result in the handle;
 * wf_get_tune/wf_set_tune move a calibration between handles.  Later solves
 * of the same WF_DEPTH_AUTO geometry at the tuned sampling wavelength use the
 * measured fastest depth; other wavelengths use the fitted model.  On a
 * geometry of fixed depth the solves keep that depth (best_depth is advisory). */
typedef struct wf_tune {
  double leaf_s_per_pair; /* leaf sum, per (point, leaf-grid direction, solve) */
  double agg_s_per_byte;  /* aggregation, per compulsory HBM byte (complex64 patterns) */
  double level_s;         /* fixed cost per level transition and solve */
  int32_t leaf_quad;      /* 1: quad-symmetric BF16x3 leaf kernel where eligible (even
                             leaf-grid n_phi), a block running several groups of 64
                             direction quads of its leaf on one staging of the leaf's
                             points where that pays; 2: the same kernel, one group per
                             block; 3: all of a leaf's groups in one block; 0: plain.
                             Results are the same for 1, 2 and 3 */
  int32_t chunk_solves;   /* solves per tree pass (1..8); opts.field_chunk overrides */
  int32_t agg_kernel;     /* level transitions: 0 = dense split-BF16 MFMA interpolation for
                             levels whose extended parent grid (2 Nt) has at most agg_fft_rows
                             rows and the LDS FFT above; 1 = the original LDS FFT kernels */
  int32_t agg_fft_rows;   /* see agg_kernel (default 600; wf_autotune lowers it while the
                             solves get faster) */
  int32_t calibrated;     /* 1 after wf_autotune (or a wf_set_tune of a calibrated tune) */
  int32_t agg_fused;      /* internal MFMA level transitions whose grids fit run as one
                             kernel, the phi output kept in LDS: 2 (default: phi matrix and
                             child rows staged in LDS, the next child loaded during the
                             current one) or 1 (phi operands read from global memory);
                             0: two passes */
  int32_t theta_tile;     /* two-pass MFMA theta pass of 6-9 K steps (161-288 great-circle
                             samples): 0 or 1 = 128 rows x 32 column pairs per block,
                             2 = 128 x 16, 3 = 64 x 16 */
  int32_t phi_rows;       /* two-pass MFMA phi pass of rows longer than 160 samples: rows
                             per staged tile, 0 or 32 (default) or 16 */
  double beta;            /* level sampling constant of the forward solves in place of
                             opts.beta (0, the default: opts.beta); set by wf_autotune.
                             wf_po_vjp always uses opts.beta */
} wf_tune;

wf_status wf_get_tune(wf_handle* h, wf_tune* out);
wf_status wf_set_tune(wf_handle* h, const wf_tune* t);

/* Times solves of `inc` (its first wavelength group, at most 8 solves) on the
 * current geometry at up to three candidate depths <= the built depth, fits
 * the cost model, picks the leaf kernel and the level-transition kernels
 * (agg_fused 2 or 0, theta_tile 0 or 2, then agg_fft_rows: the largest level
 * grids move to the LDS FFT kernels one at a time, at most three, while the
 * solves get faster) by measured time, picks the level
 * sampling constant tune.beta: the smallest of 3 and 2 below opts.beta whose
 * intensities stay within 1e-4 relative L2 (every solve of the group) of
 * the opts.beta solve, and returns the measured fastest depth in *best_depth
 * (nullable).  Synchronous. */
wf_status wf_autotune(wf_handle* h, const wf_incidence* inc, int32_t n_inc, wf_grid grid,
                      int32_t* best_depth, void* stream);

/* ---------------------------------------------------------------------------
 * Batched solves over many surfaces (no reference counterpart: the reference
 * runs one fiber segment per process, treePO.cu:21-992).  Every segment keeps
 * its own cached octree; solves run on n_streams worker streams, each with
 * its own workspace, and each (segment, stream) pair keeps its captured
 * hipGraph when opts.use_graph is set, so repeated wf_batch_solve calls with
 * the same incidences replay.  Segments are solved largest first. */
typedef struct wf_batch wf_batch;

wf_status wf_batch_create(wf_batch** out, const wf_opts* opts, int32_t n_streams);
void wf_batch_destroy(wf_batch* b);

/* Uploads a segment and builds its octree (depth in [0, 10] or WF_DEPTH_AUTO);
 * *index (nullable) receives its position in the batch output. */
wf_status wf_batch_add(wf_batch* b, const wf_surface* s, int32_t depth, int32_t* index);

/* Solves every segment for the same incidences and grid.  intensity is
 * [n_segments][n_inc][Nt][Np] float32 in out_memspace.  WF_DEPTH_AUTO
 * segments are tuned first (wf_autotune on the largest one) unless the batch
 * already holds a calibration.  WF_MEM_HOST: returns with the results in host
 * memory.  WF_MEM_DEVICE: asynchronous; `stream` is made to wait for the
 * worker streams, which first wait for `stream`'s prior work. */
wf_status wf_batch_solve(wf_batch* b, const wf_incidence* inc, int32_t n_inc, wf_grid grid,
                         float* intensity, int32_t out_memspace, void* stream);

/* Per-segment chosen depth (last wavelength group) and solve time (ms) of the
 * last wf_batch_solve; arrays of length n_segments, nullable. */
wf_status wf_batch_info(wf_batch* b, int32_t* n_segments, int32_t* depth, double* ms,
                        int64_t* n_points);

/* Finite-distance validator (compareBEM3d/ellipsoid_simple.cu:229-261 and
 * :384-589): the FP64 direct sum with the exact Green's function,
 * E = -i w mu0/(4 pi) sum_n J_n exp(-i k |x_n - ob|)/|x_n - ob| (H likewise)
 * at observation points ob = distance_m * s of the grid directions s about
 * the origin, then the far-field projection and intensity of the other
 * solves (the reference's validator uses distance 1 m).  Outputs as
 * wf_po_solve_fp64_direct. */
wf_status wf_po_solve_fp64_finite(wf_handle* h, const wf_surface* s, const wf_incidence* inc,
                                  int32_t n_inc, wf_grid grid, double distance_m,
                                  double* intensity, double* E_cplx, void* stream);

/* Geometry front-end: fiber heightfield -> surface samples
 * (util/geometry.h:78-178, set_sourcepoint with the arc window [start, end];
 * the windowless form util/geometry.h:3-72 is start = 0, end = numel - 1).
 * height: [2][numel][numelvert] float32 -- the x block, then the y block;
 * vertex (arc j, ring i) at j*numelvert + i.  Output: n = (end - start + 1) *
 * numelvert samples as 7 float32 arrays [x | y | z | nx | ny | nz | area] of
 * length n (out_soa holds 7 n floats); sample (j, i) at i*(end - start + 1) +
 * j - start is the quad (j, i), (j + 1 mod numel, i), (j, i'), (j + 1 mod
 * numel, i') with i' = min(i + 1, numelvert - 1) at heights z_i = i*unitz -
 * period/2 and z_i + unitz: centroid, mean of its two triangle areas and the
 * normalised sum of their area vectors, in the reference's float arithmetic
 * (no fused multiply-add).  A vertex with x = y = 0 is an error (the
 * reference prints it and exits); wf_last_error names the first such sample
 * in (i, j) order.  memspace: height and out_soa both host or both device;
 * the call returns after the stream has finished the conversion. */
wf_status wf_heightfield_points(const float* height, int32_t numel, int32_t numelvert, float period,
                                float unitz, int32_t start, int32_t end, float* out_soa,
                                int32_t memspace, void* stream);

/* Block average of output grids (util/fields.h:206-234, average_cpu): in
 * [n_grids][n_theta][n_phi] float32 -> out [n_grids][n_theta_out][n_phi_out]
 * with rt = n_theta / n_theta_out and rp = n_phi / n_phi_out (integer
 * quotients; rows and columns beyond rt * n_theta_out, rp * n_phi_out are
 * ignored, as there), averaging along phi first and then along theta, each
 * term divided before it is added, in float -- bit-identical to the
 * reference's loop.  memspace: in and out both host (returns with the result)
 * or both device (asynchronous on stream). */
wf_status wf_average_grid(const float* in, int32_t n_grids, int32_t n_theta, int32_t n_phi,
                          int32_t n_theta_out, int32_t n_phi_out, float* out, int32_t memspace,
                          void* stream);

/* 2D cross-section PO (compareBEM2d/rough2d.cu:237-585): a closed polyline
 * x[numel], y[numel] (metres, counter-clockwise; sample i at vertex i with the
 * length and normal (dy, -dx)/|d| of segment i -> i+1 mod numel) lit by a
 * plane wave per incidence; intensity [n_inc][n_phi] float64 at azimuths
 * phi_k = 2 pi k / n_phi in the incidence's elevation plane:
 * |E''|^2 distance_m / (2 radius_m) with the 2D far-field factor
 * sqrt(2i/(pi k_rho D)) exp(-i k_rho D), k_rho = k cos theta (the reference:
 * theta 0, phi pi, D = 1 m, radius 10 um, n_phi 3600).  FP64 throughout.
 * memspace: x, y and intensity all host (returns with the result) or all
 * device (the call still returns after the stream has finished, to report
 * a zero-length segment, which is an error). */
typedef struct wf_rough2d_incidence {
  double theta;          /* elevation of incidence and observation (rough2d.cu:464) */
  double phi;            /* incidence azimuth (rough2d.cu:471: pi) */
  double lambda_m;       /* vacuum wavelength (rough2d.cu:467-468) */
  double eta_re, eta_im; /* complex IOR, outside medium 1 (rough2d.cu:466) */
  char pol;              /* 'M' = TM (E along z at theta 0), 'E' = TE (rough2d.cu:269-277) */
} wf_rough2d_incidence;

wf_status wf_rough2d_solve(const double* x, const double* y, int32_t numel,
                           const wf_rough2d_incidence* inc, int32_t n_inc, int32_t n_phi,
                           double distance_m, double radius_m, double* intensity,
                           int32_t memspace, void* stream);

/* Allocation-free variant of wf_rough2d_solve.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * hipGraph capture, tight serving loops, harnesses that forbid allocation).
 * Same arguments and results as wf_rough2d_solve with memspace
 * WF_MEM_DEVICE (x, y and intensity are device pointers; inc is host memory),
 * and the call likewise returns after the stream has finished.  Its scratch
 * memory -- the incidence parameters, the currents (96 B per sample and
 * incidence) and the partial sums of the azimuth kernel (96 B per split,
 * incidence and azimuth) -- is one static device array of
 * WF_ROUGH2D_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends;
 * larger problems, or more than WF_ROUGH2D_NOALLOC_MAX_INC incidences, are
 * rejected with WF_ERR_INVALID_ARG.  Every call on a GPU shares that array
 * and it keeps the last call's contents, so calls that overlapped on one GPU
 * would race on it: the library serialises the calls of a process (each
 * holds a lock until its stream has finished), so concurrent callers wait. */
#define WF_ROUGH2D_NOALLOC_SCRATCH_BYTES (64 << 20)
#define WF_ROUGH2D_NOALLOC_MAX_INC 1024
wf_status wf_rough2d_solve_noalloc(const double* x, const double* y, int32_t numel,
                                   const wf_rough2d_incidence* inc, int32_t n_inc, int32_t n_phi,
                                   double distance_m, double radius_m, double* intensity,
                                   void* stream);

/* Reads a fiber heightfield file (treePO.cu:79-88: raw float32, 2 * numel *
 * numelvert values in host byte order) into host memory; a missing or short
 * file is an error. */
wf_status wf_read_heightfield(const char* path, int32_t numel, int32_t numelvert, float* out);

#ifdef __cplusplus
}
#endif
#endif /* WF_PO_H */
