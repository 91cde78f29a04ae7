/*
 * fsoracle.h -- f64 CPU oracle for the stable-fluids step on a periodic MAC grid.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP product
 * path is compared against; it is never linked into, loaded by, or called from
 * the product (fluid-sim_amd/). Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it.
 *
 * It restates, in plain C and double precision, the reference's algorithm:
 *   src/core/grid.js      MAC layout, bilerp, samplers, div/grad, diagnostics
 *   src/core/analytic.js  Taylor-Green closed form, error norms
 *   src/core/fft.js       radix-2 FFT, fft2/ifft2 normalisation, energy spectrum
 *   src/cpu/advect.js     RK2 trace, semi-Lagrangian and limited MacCormack
 *   src/cpu/projection.js Laplacian, CG, FFT Poisson solve
 *   src/cpu/solver.js     explicit/implicit diffusion, band forcing, drag, step order
 *   src/gpu/shaders.js + src/gpu/multigrid.js
 *                         Jacobi / red-black / V-cycle semantics of the GPU
 *                         solver, restated in f64 so the pressure solvers have
 *                         an oracle too.
 * Each function in fsoracle.c cites the reference file:line it follows.
 *
 * Layout: every field is nx*ny doubles, row-major, idx = j*nx + i.
 *   u[i,j] at (i*dx, (j+0.5)*dy)   left face
 *   v[i,j] at ((i+0.5)*dx, j*dy)   bottom face
 *   p[i,j] at ((i+0.5)*dx, (j+0.5)*dy) centre
 */
#ifndef FSORACLE_H
#define FSORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- projection / advection selectors for or_step ---- */
enum { OR_ADV_SL = 0, OR_ADV_MC = 1 };
enum { OR_PROJ_FFT = 0, OR_PROJ_CG = 1, OR_PROJ_JACOBI = 2, OR_PROJ_REDBLACK = 3,
       OR_PROJ_MULTIGRID = 4, OR_PROJ_NONE = 5 };

typedef struct {
  int nx, ny;
  double lx, ly;
  double nu;
  int advection;          /* OR_ADV_* */
  int explicit_diffusion; /* 1: explicit 5-point, 0: implicit CG (reference CPU default) */
  int projection;         /* OR_PROJ_* */
  int iterations;         /* sweeps (jacobi, red-black) or V-cycles (multigrid) */
  int mg_min_size, mg_pre, mg_post, mg_coarse;
  double cg_tol;
  int cg_maxit;
  /* band forcing (src/cpu/solver.js:179-237); off when forcing == 0 */
  int forcing;
  double kf, width, amp, drag;
  /* one immersed disk with prescribed rigid motion (0 = none) */
  int solid;
  double sx, sy, sr, svx, svy, somega, mask_width;
  int passes;
  /* the GPU's mode-list forcing (src/gpu/solver.js:370-414) instead of the band forcing
   * when modes > 0: kx, ky, phase of each mode, amplitude, drag, phase drift, and the
   * simulated time t at the start of the step */
  int modes;
  const double *mkx, *mky, *mph;
  double mamp, mdrag, mdrift, mt;
} or_step_cfg;

/* rigid disk for the direct-forcing coupling (src/cpu/solid.js:52-114) */
typedef struct {
  double x, y, r, vx, vy, omega;
} or_disk;

void   or_set_threads(int n);
int    or_get_threads(void);

/* seeded generator of the reference tests: s=(s*1103515245+12345)&0x7fffffff in IEEE double */
void   or_js_lcg_fill(int64_t *state, int n, double *out);

int    or_wrap(int i, int n);
double or_bilerp(const double *a, int nx, int ny, double gx, double gy);
double or_sample_u(int nx, int ny, double lx, double ly, const double *u, double x, double y);
double or_sample_v(int nx, int ny, double lx, double ly, const double *v, double x, double y);
double or_sample_p(int nx, int ny, double lx, double ly, const double *p, double x, double y);

void   or_divergence(int nx, int ny, double lx, double ly, const double *u, const double *v, double *out);
void   or_subtract_gradient(int nx, int ny, double lx, double ly, double *u, double *v,
                            const double *p, double scale);
void   or_laplacian(int nx, int ny, double lx, double ly, const double *p, double *out);
void   or_vorticity_corner(int nx, int ny, double lx, double ly, const double *u, const double *v, double *out);
void   or_velocity_at_centers(int nx, int ny, const double *u, const double *v, double *cu, double *cv);

double or_kinetic_energy(int nx, int ny, const double *u, const double *v);
double or_kinetic_energy_faces(int n, const double *u, const double *v);
double or_enstrophy(int nx, int ny, double lx, double ly, const double *u, const double *v);
double or_max_divergence(int nx, int ny, double lx, double ly, const double *u, const double *v);
double or_cfl(int nx, int ny, double lx, double ly, const double *u, const double *v, double dt);
double or_poisson_residual_rel(int nx, int ny, double lx, double ly, const double *p, const double *rhs);

void   or_advect_velocity_sl(int nx, int ny, double lx, double ly, const double *u, const double *v,
                             double *ou, double *ov, double dt);
void   or_advect_velocity_mc(int nx, int ny, double lx, double ly, const double *u, const double *v,
                             double *ou, double *ov, double dt);
void   or_advect_scalar_sl(int nx, int ny, double lx, double ly, const double *u, const double *v,
                           const double *q, double *out, double dt);
void   or_advect_scalar_mc(int nx, int ny, double lx, double ly, const double *u, const double *v,
                           const double *q, double *out, double dt);

void   or_diffuse_explicit(int nx, int ny, double lx, double ly, const double *f, double *out,
                           double nu, double dt);
int    or_diffuse_implicit(int nx, int ny, double lx, double ly, const double *f, double *out,
                           double nu, double dt, double tol);

int    or_fft(int n, double *re, double *im, int sign);
int    or_fft2(int nx, int ny, double *re, double *im, int sign);
int    or_poisson_fft(int nx, int ny, double lx, double ly, const double *rhs, double *p);
int    or_poisson_cg(int nx, int ny, double lx, double ly, const double *rhs, double *p,
                     double tol, int maxit, double *residual_out);

/* GPU-solver semantics (src/gpu/shaders.js, src/gpu/multigrid.js), in f64 */
void   or_jacobi(int nx, int ny, double lx, double ly, const double *rhs, double *p, int iters, double omega);
void   or_redblack(int nx, int ny, double lx, double ly, const double *rhs, double *p, int sweeps, double omega);
int    or_mg_levels(int n, int min_size);
int    or_mg_vcycles(int n, double lx, const double *rhs, double *p, int cycles,
                     int min_size, int pre, int post, int coarse, double *history);

/* forcing (src/cpu/solver.js:179-237) */
int    or_band_modes(uint32_t seed, double kf, double width, int *kx, int *ky, double *phase);
void   or_band_forcing(int nx, int ny, double lx, double ly, double *u, double *v,
                       double kf, double width, double amp, uint32_t seed, double dt);
void   or_linear_drag(int n, double *u, double *v, double alpha, double dt);

/* closed forms (src/core/analytic.js) */
void   or_taylor_green(int nx, int ny, double lx, double ly, double t, double nu, double *u, double *v);
double or_velocity_error_tg(int nx, int ny, double lx, double ly, const double *u, const double *v,
                            double t, double nu);

/* shell-binned energy spectrum (src/core/fft.js:196-234); E has kmax+1 entries */
int    or_energy_spectrum(int nx, int ny, const double *cu, const double *cv, double *E, int *counts);

/* storage rounding (0: FP32, 1: BF16) and the refinement projection of the GPU
 * solver with the correction stored at that precision */
double or_round_storage(double x, int storage);
int    or_mg_refine(int n, double lx, const double *div, double *p, int iterations, int min_size, int pre,
                    int post, int coarse, int coarse_n, int storage, int coarse_exact, double *rmax);
int    or_poisson_dft(int n, double lx, const double *rhs, double *x);

/* immersed disk (src/cpu/solid.js:38-40, :101-114, :165-233) */
double or_wrap_delta(double d, double l);
double or_disk_mask(const or_disk *s, double px, double py, double lx, double ly, double width);
void   or_disk_velocity(const or_disk *s, double px, double py, double lx, double ly, double *vx, double *vy);
/* `passes` direct-forcing passes in place; imp = fluid impulse (x, y) and torque impulse, dA included */
void   or_apply_coupling(int nx, int ny, double lx, double ly, double *u, double *v, const or_disk *s,
                         double mask_width, int passes, double *imp);

/* lid-driven cavity on the unit box (walls; no reference counterpart) */
double or_cav_sample_u(int n, double U, const double *u, double x, double y);
double or_cav_sample_v(int n, const double *v, double x, double y);
void   or_cav_advect(int n, double U, int scheme, const double *u, const double *v, double *ou, double *ov,
                     double dt);
void   or_cav_diffuse(int n, double U, const double *u, const double *v, double *ou, double *ov, double nu,
                      double dt);
void   or_cav_divergence(int n, const double *u, const double *v, double *div);
void   or_cav_laplacian(int n, const double *p, double *out);
void   or_cav_subtract_gradient(int n, double *u, double *v, const double *p);
int    or_cav_poisson_exact(int n, const double *rhs, double *p);
void   or_cav_apply_coupling(int n, double *u, double *v, const or_disk *s, double mask_width, int passes,
                             double *imp);
int    or_cav_step(int n, double U, double nu, int scheme, const or_disk *disk, double mask_width, int passes,
                   double dt, double *u, double *v, double *p_out);

/* mode-list forcing of the GPU path (FORCING src/gpu/shaders.js:532-566, parameters
 * src/gpu/solver.js:387-396): u <- u exp(-drag dt) + amp dt / sqrt(count) *
 * sum (ky/|k|) cos(kx x + ky y + phase + drift t) at the u faces, -(kx/|k|) at the v
 * faces; modes with |k| < 1e-6 add nothing but count */
void   or_mode_forcing(int nx, int ny, double lx, double ly, double *u, double *v, int count, const double *kx,
                       const double *ky, const double *phase, double amp, double drag, double drift, double t,
                       double dt);

/* the GPU's dye step (ADVECT_SCALAR src/gpu/shaders.js:234-266): each of the four channels
 * of the cell-centred rgba dye (dye[4 k + c]) advected semi-Lagrangian through (u, v),
 * times (1 - fade), clamped to [0, 1] */
void   or_advect_dye(int nx, int ny, double lx, double ly, const double *u, const double *v, const double *dye,
                     double *out, double dt, double fade);

/* one splat (SPLAT src/gpu/shaders.js:474-511): g = exp(-|d|^2 / max(radius^2, 1e-12)) with
 * d the periodic shortest separation (round half to even, as WGSL's round) of the cell
 * centre from (x, y); a dye splat adds (r, g, b, 0) g to the rgba dye, a force splat
 * adds (dx g, dy g) to the u and v faces of the cell */
typedef struct {
  double x, y, dx, dy, r, g, b, radius;
  int is_dye;
} or_splat;
void   or_splat_apply(int nx, int ny, double lx, double ly, double *u, double *v, double *dye, const or_splat *s);

/* one full step (src/cpu/solver.js:336-398); p_out may be NULL */
int    or_step(const or_step_cfg *cfg, double *u, double *v, double *p_out, double dt, uint32_t force_seed);

#ifdef __cplusplus
}
#endif
#endif
