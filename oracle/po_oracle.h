/*
 * po_oracle.h -- FP64 CPU restatement of WaveFiber3d's physical-optics (PO)
 * far-field path.  TEST INFRASTRUCTURE ONLY: the product library never links,
 * loads or calls this code.  It is the checker for the HIP kernels (tests,
 * the smoke check and the CPU-baseline leg of bench.py).
 *
 * Every function cites the reference file:line it restates.  All arithmetic is
 * double / complex double; inputs arrive as float32 (the reference's storage
 * type) and are widened exactly.
 */
#ifndef WF_PO_ORACLE_H
#define WF_PO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants, util/util.h:16-19 (float there, double here). */
#define ORC_C0 299792458.0
#define ORC_MU0 (4e-7 * 3.14159265358979323846)
#define ORC_EPS0 8.8541878e-12

typedef struct {
  double theta;        /* incidence latitude (rad), treePO.cu:560 convention */
  double phi;          /* incidence azimuth (rad) */
  double lambda_m;     /* vacuum wavelength (m) */
  char pol;            /* 'M' (TM) or 'E' (TE), treePO.cu:573-601 */
  double eta_re, eta_im; /* complex IOR of the body (outside medium is 1) */
  double beam_sigma_m; /* Gaussian taper sigma on z (m); <= 0 or inf: off */
} orc_incidence;

double orc_z0(void);

/* util/util.h:365-382 with n0 = 1: rs, rp at local angle theta. out = {re, im}. */
void orc_fresnel(double theta_loc, double eta_re, double eta_im, double rs[2], double rp[2]);

/* fibersim/treePO.cu:580-601: propagation dir w, E and H unit vectors. */
void orc_incidence_basis(double theta, double phi, char pol, double w[3], double E[3], double H[3]);

/* fibersim/treePO.cu:504-509: E_f, H_f = -i w {mu0, eps0}/(4 pi) e^{-i k D}/D, D = 1 m. */
void orc_radiation_factors(double lambda_m, double Ef[2], double Hf[2]);

/* util/tangentplane.h:18-80 + util/fields.h:106-150: scaled currents
 * C[n][6][2] = a_n * [E_f J ; H_f M]  (J = n x H_tot, M = E_tot x n). */
void orc_currents(int64_t n, const float* x, const float* y, const float* z,
                  const float* nx, const float* ny, const float* nz, const float* area,
                  const orc_incidence* inc, double* C);

/* fibersim/treePO.cu:378-391: theta_j = -pi/2 + 1e-10 + pi j/Nt, phi_k = 2 pi k/Np,
 * s = (cos t cos p, cos t sin p, sin t), flat index j*Np + k. dirs[D][3]. */
void orc_output_dirs(int ntheta, int nphi, double* dirs);

/* util/fields.h:8-38 + treePO.cu:662-670 at depth 0: for S solves sharing k,
 * F[s][d][c] = sum_n C[s][n][c] exp(+i k s_d.(x_n - centre)).  OpenMP over d. */
void orc_direct_sum(int64_t n, const float* x, const float* y, const float* z,
                    const double centre[3], double k, int S, const double* C,
                    int64_t ndir, const double* dirs, double* F, int nthreads);

/* compareBEM3d/ellipsoid_simple.cu:229-261, :547-555: the sum above with the
 * exact Green's function exp(-i k |x - ob|)/|x - ob| at ob = dist * s_d about
 * the origin, times D0 exp(+i k D0) (D0 = 1 m) to remove the far-field factor
 * the currents carry. */
void orc_direct_sum_finite(int64_t n, const float* x, const float* y, const float* z, double k,
                           int S, const double* C, int64_t ndir, const double* dirs, double dist,
                           double* F, int nthreads);

/* util/fields.h:172-204: E'' = E_t - Z0 s x H_t, I = |E''|^2.
 * F[d][6][2] -> I[d], and (optional) E[3][d][2]. */
void orc_intensity(int64_t ndir, const double* dirs, const double* F, double* I, double* E);

/* Full depth-0 solve (the oracle semantics, sphere.cu:592-658 + :909-923 with
 * depth 0): I[i][Nt*Np]; E (nullable) [i][3][Nt*Np][2].  Returns 0 on success. */
int orc_solve_direct(int64_t n, const float* x, const float* y, const float* z,
                     const float* nx, const float* ny, const float* nz, const float* area,
                     const orc_incidence* inc, int n_inc, int ntheta, int nphi,
                     const double centre[3], double* I, double* E, int nthreads);

/* Same as orc_solve_direct but on an explicit direction list dirs[D][3]. */
int orc_solve_direct_dirs(int64_t n, const float* x, const float* y, const float* z,
                          const float* nx, const float* ny, const float* nz, const float* area,
                          const orc_incidence* inc, int n_inc, int64_t ndir, const double* dirs,
                          const double centre[3], double* I, double* E, int nthreads);

/* orc_solve_direct_dirs with orc_direct_sum_finite at distance dist
 * (ellipsoid_simple.cu:384-589 semantics). */
int orc_solve_finite_dirs(int64_t n, const float* x, const float* y, const float* z,
                          const float* nx, const float* ny, const float* nz, const float* area,
                          const orc_incidence* inc, int n_inc, int64_t ndir, const double* dirs,
                          double dist, double* Iout, double* E, int nthreads);

/* util/tree.h:289-715 (+ treePO.cu:219-259): full octree to `depth` over the
 * float box {min x,y,z, max x,y,z}, recursive stable 8-way partition with the
 * reference's float centre and octant predicates.  Outputs (node arrays hold
 * (8^(depth+1)-1)/7 entries in level order):
 *   perm[i]       original index of the point at sorted position i
 *   node_begin/end  point range per node
 *   node_box[6*j] float box per node. */
int orc_octree(int64_t n, const float* x, const float* y, const float* z,
               const float box[6], int depth, int64_t* perm,
               int32_t* node_begin, int32_t* node_end, float* node_box);

/* treePO.cu:342-375 / sphere.cu:345-365 sampling rule for level i >= 1:
 * N = ceil(k d + beta (k d)^(1/3)), k = 2 pi/0.4e-6, d = 0.5^i * extent. */
int orc_ref_level_count(int level, float extent, float beta);

/* util/interpolation.h:11-80 semantics on one line, evaluated by brute force:
 * the trigonometric interpolant of nc samples taken at t_j = off_c + 2 pi j/nc
 * (Nyquist bin split in half when nc is even) evaluated at
 * u_m = off_p + 2 pi m/np.  in/out complex interleaved. */
void orc_trig_interp(int nc, const double* in, double off_c, int np, double off_p, double* out);

/* Sphere-pattern interpolation used by the tree (2 pi theta extension):
 * child pattern on the grid theta_j = off_tc + pi j/ntc, phi_k = 2 pi k/npc
 * (ntc x npc, npc even) -> values at the parent grid theta_j = off_tp + pi j/ntp,
 * phi_k = 2 pi k/npp.  Brute force via orc_trig_interp along phi, then along the
 * great circle formed by columns phi and phi + pi. */
/* util/geometry.h:78-178 heightfield -> samples (float arithmetic, no FMA);
 * 0 or 1 + i*numel + j of the first sample with a zero vertex. */
long orc_heightfield_points(const float* h, int numel, int numelvert, float period, float unitz,
                            int start, int end, float* x, float* y, float* z, float* nx,
                            float* ny, float* nz, float* area);
/* util/fields.h:206-234 average_cpu on one grid [nt][np] -> [nto][npo]. */
void orc_average_grid(const float* in, int nt, int np, int nto, int npo, float* out);
void orc_sphere_interp(int ntc, int npc, double off_tc, const double* in,
                       int ntp, int npp, double off_tp, double* out);

/* compareBEM2d/rough2d.cu:237-585 for one incidence: closed polyline x, y [n]
 * (sample i at vertex i, segment i -> i+1 mod n), scattering [nphi] =
 * |E''|^2 Dis/(2 radius) at phi_p = 2 pi p/nphi; C (nullable) receives the
 * unscaled currents [n][6][2], components Jx, Jy, Jz, Mx, My, Mz (rough2d.cu:348-361). */
void orc_rough2d(int n, const double* x, const double* y, double theta, double phii,
                 double lambda_m, char mode, double eta_re, double eta_im, int nphi, double Dis,
                 double radius, double* scattering, double* C);

#ifdef __cplusplus
}
#endif
#endif
