/*
 * fsoracle.c -- f64 CPU oracle for the stable-fluids step (TEST INFRASTRUCTURE).
 *
 * This is the checker, not the product: the HIP library in fluid-sim_amd/ never
 * links or loads it. Each function names the reference file:line whose
 * algorithm it restates; the restatement is written from the mathematics the
 * reference documents, in C, with the same operation order where that order
 * decides the result (sampling offsets, limiter bounds, colour order).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   LCG, divergence       reference-held max|div| of the seed-4242 field
 *   Jacobi / red-black /  reference-held residual curves of the seed-777 field
 *   multigrid             (results/gpu-verify.json poissonCurves)
 *   SL / MacCormack       reference-held energies after 200 TG steps
 *                         (results/gpu-verify.json schemeEnergy), integer-shift
 *                         exactness, translating-Gaussian order
 *   FFT / Poisson         naive DFT, single-mode known answers, Parseval
 *   diffusion             closed-form checkerboard and sine-mode factors
 *   forcing               exact integer LCG, curl structure
 */
#include "fsoracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

void or_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int or_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* test/core.test.js:94-98 -- the product s*1103515245 exceeds 2^53, so it is
 * rounded to double before ToInt32 takes it mod 2^32; fmod on doubles is exact. */
void or_js_lcg_fill(int64_t *state, int n, double *out) {
  int64_t s = *state;
  for (int k = 0; k < n; k++) {
    double prod = (double)s * 1103515245.0;
    double sum = prod + 12345.0;
    double m = fmod(sum, 4294967296.0);
    if (m < 0) m += 4294967296.0;
    uint32_t u32 = (uint32_t)m;
    s = (int64_t)(u32 & 0x7fffffffu);
    out[k] = (double)s / 2147483647.0 - 0.5;
  }
  *state = s;
}

/* src/core/grid.js:40-42 */
int or_wrap(int i, int n) { return ((i % n) + n) % n; }

/* src/core/grid.js:147-165 */
double or_bilerp(const double *a, int nx, int ny, double gx, double gy) {
  double fi = floor(gx), fj = floor(gy);
  int i0 = (int)fi, j0 = (int)fj;
  double fx = gx - fi, fy = gy - fj;
  int i0w = or_wrap(i0, nx), i1w = or_wrap(i0 + 1, nx);
  int j0w = or_wrap(j0, ny), j1w = or_wrap(j0 + 1, ny);
  double A = a[j0w * nx + i0w], B = a[j0w * nx + i1w];
  double C = a[j1w * nx + i0w], D = a[j1w * nx + i1w];
  double top = A + (B - A) * fx;
  double bot = C + (D - C) * fx;
  return top + (bot - top) * fy;
}

/* src/core/grid.js:167-180 */
double or_sample_u(int nx, int ny, double lx, double ly, const double *u, double x, double y) {
  double dx = lx / nx, dy = ly / ny;
  return or_bilerp(u, nx, ny, x / dx, y / dy - 0.5);
}
double or_sample_v(int nx, int ny, double lx, double ly, const double *v, double x, double y) {
  double dx = lx / nx, dy = ly / ny;
  return or_bilerp(v, nx, ny, x / dx - 0.5, y / dy);
}
double or_sample_p(int nx, int ny, double lx, double ly, const double *p, double x, double y) {
  double dx = lx / nx, dy = ly / ny;
  return or_bilerp(p, nx, ny, x / dx - 0.5, y / dy - 0.5);
}

/* src/core/grid.js:191-206 */
void or_divergence(int nx, int ny, double lx, double ly, const double *u, const double *v, double *out) {
  double dx = lx / nx, dy = ly / ny;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    int jp = (j + 1 == ny) ? 0 : j + 1;
    for (int i = 0; i < nx; i++) {
      int ip = (i + 1 == nx) ? 0 : i + 1;
      double uL = u[j * nx + i], uR = u[j * nx + ip];
      double vB = v[j * nx + i], vT = v[jp * nx + i];
      out[j * nx + i] = (uR - uL) / dx + (vT - vB) / dy;
    }
  }
}

/* src/core/grid.js:209-218 */
void or_subtract_gradient(int nx, int ny, double lx, double ly, double *u, double *v,
                          const double *p, double scale) {
  double dx = lx / nx, dy = ly / ny;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    int jm = (j == 0) ? ny - 1 : j - 1;
    for (int i = 0; i < nx; i++) {
      int im = (i == 0) ? nx - 1 : i - 1;
      u[j * nx + i] -= scale * (p[j * nx + i] - p[j * nx + im]) / dx;
      v[j * nx + i] -= scale * (p[j * nx + i] - p[jm * nx + i]) / dy;
    }
  }
}

/* src/cpu/projection.js:48-62 */
void or_laplacian(int nx, int ny, double lx, double ly, const double *p, double *out) {
  double dx = lx / nx, dy = ly / ny;
  double ix = 1.0 / (dx * dx), iy = 1.0 / (dy * dy);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    int jm = or_wrap(j - 1, ny) * nx, jp = or_wrap(j + 1, ny) * nx, j0 = j * nx;
    for (int i = 0; i < nx; i++) {
      int im = or_wrap(i - 1, nx), ip = or_wrap(i + 1, nx);
      out[j0 + i] = (p[j0 + im] - 2 * p[j0 + i] + p[j0 + ip]) * ix +
                    (p[jm + i] - 2 * p[j0 + i] + p[jp + i]) * iy;
    }
  }
}

/* src/core/grid.js:229-240 */
void or_vorticity_corner(int nx, int ny, double lx, double ly, const double *u, const double *v, double *out) {
  double dx = lx / nx, dy = ly / ny;
  for (int j = 0; j < ny; j++) {
    int jm = or_wrap(j - 1, ny);
    for (int i = 0; i < nx; i++) {
      int im = or_wrap(i - 1, nx);
      double dvdx = (v[j * nx + i] - v[j * nx + im]) / dx;
      double dudy = (u[j * nx + i] - u[jm * nx + i]) / dy;
      out[j * nx + i] = dvdx - dudy;
    }
  }
}

/* src/core/grid.js:267-279 */
void or_velocity_at_centers(int nx, int ny, const double *u, const double *v, double *cu, double *cv) {
  for (int j = 0; j < ny; j++) {
    int jp = or_wrap(j + 1, ny);
    for (int i = 0; i < nx; i++) {
      int ip = or_wrap(i + 1, nx);
      cu[j * nx + i] = 0.5 * (u[j * nx + i] + u[j * nx + ip]);
      cv[j * nx + i] = 0.5 * (v[j * nx + i] + v[jp * nx + i]);
    }
  }
}

/* src/core/grid.js:290-295 (centre-interpolated) */
double or_kinetic_energy(int nx, int ny, const double *u, const double *v) {
  double s = 0;
  for (int j = 0; j < ny; j++) {
    int jp = or_wrap(j + 1, ny);
    for (int i = 0; i < nx; i++) {
      int ip = or_wrap(i + 1, nx);
      double cu = 0.5 * (u[j * nx + i] + u[j * nx + ip]);
      double cv = 0.5 * (v[j * nx + i] + v[jp * nx + i]);
      s += cu * cu + cv * cv;
    }
  }
  return 0.5 * s / ((double)nx * ny);
}

/* tools/gputest.html:313-316 -- raw face values, a different definition */
double or_kinetic_energy_faces(int n, const double *u, const double *v) {
  double s = 0;
  for (int k = 0; k < n; k++) s += u[k] * u[k] + v[k] * v[k];
  return 0.5 * s / n;
}

/* src/core/grid.js:298-303 with vorticityCenter :245-262 */
double or_enstrophy(int nx, int ny, double lx, double ly, const double *u, const double *v) {
  double *w = (double *)malloc(sizeof(double) * nx * ny);
  or_vorticity_corner(nx, ny, lx, ly, u, v, w);
  double s = 0;
  for (int j = 0; j < ny; j++) {
    int jp = or_wrap(j + 1, ny);
    for (int i = 0; i < nx; i++) {
      int ip = or_wrap(i + 1, nx);
      double c = 0.25 * (w[j * nx + i] + w[j * nx + ip] + w[jp * nx + i] + w[jp * nx + ip]);
      s += c * c;
    }
  }
  free(w);
  return 0.5 * s / ((double)nx * ny);
}

/* src/core/grid.js:306-311 */
double or_max_divergence(int nx, int ny, double lx, double ly, const double *u, const double *v) {
  double *d = (double *)malloc(sizeof(double) * nx * ny);
  or_divergence(nx, ny, lx, ly, u, v, d);
  double m = 0;
  for (int k = 0; k < nx * ny; k++) m = fmax(m, fabs(d[k]));
  free(d);
  return m;
}

/* src/core/grid.js:322-327 */
double or_cfl(int nx, int ny, double lx, double ly, const double *u, const double *v, double dt) {
  double mu = 0, mv = 0;
  for (int k = 0; k < nx * ny; k++) { mu = fmax(mu, fabs(u[k])); mv = fmax(mv, fabs(v[k])); }
  return dt * (mu / (lx / nx) + mv / (ly / ny));
}

/* src/cpu/projection.js:251-261 */
double or_poisson_residual_rel(int nx, int ny, double lx, double ly, const double *p, const double *rhs) {
  double *Ap = (double *)malloc(sizeof(double) * nx * ny);
  or_laplacian(nx, ny, lx, ly, p, Ap);
  double num = 0, den = 0;
  for (int k = 0; k < nx * ny; k++) {
    double d = Ap[k] - rhs[k];
    num += d * d;
    den += rhs[k] * rhs[k];
  }
  free(Ap);
  return sqrt(num) / (den > 0 ? sqrt(den) : 1.0);
}

/* ------------------------------------------------------------------ advection */

typedef struct { int nx, ny; double dx, dy; } G;

static inline double su(const G *g, const double *u, double x, double y) {
  return or_bilerp(u, g->nx, g->ny, x / g->dx, y / g->dy - 0.5);
}
static inline double sv(const G *g, const double *v, double x, double y) {
  return or_bilerp(v, g->nx, g->ny, x / g->dx - 0.5, y / g->dy);
}

/* src/cpu/advect.js:46-54 -- RK2 midpoint backward trace */
static inline void trace_back(const G *g, const double *u, const double *v, double x, double y,
                              double dt, double *ox, double *oy) {
  double u1 = su(g, u, x, y), v1 = sv(g, v, x, y);
  double xm = x - 0.5 * dt * u1, ym = y - 0.5 * dt * v1;
  double u2 = su(g, u, xm, ym), v2 = sv(g, v, xm, ym);
  *ox = x - dt * u2;
  *oy = y - dt * v2;
}

/* src/cpu/advect.js:74-97 */
void or_advect_velocity_sl(int nx, int ny, double lx, double ly, const double *u, const double *v,
                           double *ou, double *ov, double dt) {
  G g = {nx, ny, lx / nx, ly / ny};
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      double x, y;
      trace_back(&g, u, v, i * g.dx, (j + 0.5) * g.dy, dt, &x, &y);
      ou[j * nx + i] = su(&g, u, x, y);
      trace_back(&g, u, v, (i + 0.5) * g.dx, j * g.dy, dt, &x, &y);
      ov[j * nx + i] = sv(&g, v, x, y);
    }
  }
}

/* src/cpu/advect.js:113-121 */
static inline void interp_bounds(const double *a, int nx, int ny, double gx, double gy,
                                 double *lo, double *hi) {
  int i0 = (int)floor(gx), j0 = (int)floor(gy);
  int i0w = or_wrap(i0, nx), i1w = or_wrap(i0 + 1, nx);
  int j0w = or_wrap(j0, ny), j1w = or_wrap(j0 + 1, ny);
  double A = a[j0w * nx + i0w], B = a[j0w * nx + i1w];
  double C = a[j1w * nx + i0w], D = a[j1w * nx + i1w];
  *lo = fmin(fmin(A, B), fmin(C, D));
  *hi = fmax(fmax(A, B), fmax(C, D));
}

static inline double clampd(double c, double lo, double hi) { return c < lo ? lo : (c > hi ? hi : c); }

/* src/cpu/advect.js:149-185 with advectComponentWith :192-206 */
void or_advect_velocity_mc(int nx, int ny, double lx, double ly, const double *u, const double *v,
                           double *ou, double *ov, double dt) {
  G g = {nx, ny, lx / nx, ly / ny};
  size_t n = (size_t)nx * ny;
  double *fu = (double *)malloc(sizeof(double) * n), *fv = (double *)malloc(sizeof(double) * n);
  or_advect_velocity_sl(nx, ny, lx, ly, u, v, fu, fv, dt);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      int k = j * nx + i;
      double x, y, lo, hi;
      /* u face: backward pass of the forward result through the ORIGINAL flow */
      double px = i * g.dx, py = (j + 0.5) * g.dy;
      trace_back(&g, u, v, px, py, -dt, &x, &y);
      double bu = or_bilerp(fu, nx, ny, x / g.dx, y / g.dy - 0.5);
      double cu = fu[k] + 0.5 * (u[k] - bu);
      trace_back(&g, u, v, px, py, dt, &x, &y);
      interp_bounds(u, nx, ny, x / g.dx, y / g.dy - 0.5, &lo, &hi);
      ou[k] = clampd(cu, lo, hi);
      /* v face */
      px = (i + 0.5) * g.dx; py = j * g.dy;
      trace_back(&g, u, v, px, py, -dt, &x, &y);
      double bv = or_bilerp(fv, nx, ny, x / g.dx - 0.5, y / g.dy);
      double cv = fv[k] + 0.5 * (v[k] - bv);
      trace_back(&g, u, v, px, py, dt, &x, &y);
      interp_bounds(v, nx, ny, x / g.dx - 0.5, y / g.dy, &lo, &hi);
      ov[k] = clampd(cv, lo, hi);
    }
  }
  free(fu); free(fv);
}

/* src/cpu/advect.js:61-71 */
void or_advect_scalar_sl(int nx, int ny, double lx, double ly, const double *u, const double *v,
                         const double *q, double *out, double dt) {
  G g = {nx, ny, lx / nx, ly / ny};
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      double x, y;
      trace_back(&g, u, v, (i + 0.5) * g.dx, (j + 0.5) * g.dy, dt, &x, &y);
      out[j * nx + i] = or_bilerp(q, nx, ny, x / g.dx - 0.5, y / g.dy - 0.5);
    }
  }
}

/* src/cpu/advect.js:124-146 */
void or_advect_scalar_mc(int nx, int ny, double lx, double ly, const double *u, const double *v,
                         const double *q, double *out, double dt) {
  G g = {nx, ny, lx / nx, ly / ny};
  size_t n = (size_t)nx * ny;
  double *fwd = (double *)malloc(sizeof(double) * n), *back = (double *)malloc(sizeof(double) * n);
  or_advect_scalar_sl(nx, ny, lx, ly, u, v, q, fwd, dt);
  or_advect_scalar_sl(nx, ny, lx, ly, u, v, fwd, back, -dt);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      int k = j * nx + i;
      double corrected = fwd[k] + 0.5 * (q[k] - back[k]);
      double x, y, lo, hi;
      trace_back(&g, u, v, (i + 0.5) * g.dx, (j + 0.5) * g.dy, dt, &x, &y);
      interp_bounds(q, nx, ny, x / g.dx - 0.5, y / g.dy - 0.5, &lo, &hi);
      out[k] = clampd(corrected, lo, hi);
    }
  }
  free(fwd); free(back);
}

/* ADVECT_SCALAR, src/gpu/shaders.js:234-266: traceBack from the cell centre, sampleC of
 * each channel, fade, clamp to [0, 1] */
void or_advect_dye(int nx, int ny, double lx, double ly, const double *u, const double *v, const double *dye,
                   double *out, double dt, double fade) {
  G g = {nx, ny, lx / nx, ly / ny};
  size_t n = (size_t)nx * ny;
  double *ch = (double *)malloc(sizeof(double) * n), *o = (double *)malloc(sizeof(double) * n);
  for (int c = 0; c < 4; c++) {
    for (size_t k = 0; k < n; k++) ch[k] = dye[4 * k + c];
    or_advect_scalar_sl(nx, ny, lx, ly, u, v, ch, o, dt);
    for (size_t k = 0; k < n; k++) {
      double q = o[k] * (1.0 - fade);
      out[4 * k + c] = q < 0 ? 0 : (q > 1 ? 1 : q);
    }
  }
  (void)g;
  free(ch);
  free(o);
}

/* SPLAT, src/gpu/shaders.js:474-511 */
void or_splat_apply(int nx, int ny, double lx, double ly, double *u, double *v, double *dye, const or_splat *sp) {
  double dx = lx / nx, dy = ly / ny;
  double r2 = sp->radius * sp->radius;
  if (r2 < 1e-12) r2 = 1e-12;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double px = (i + 0.5) * dx - sp->x, py = (j + 0.5) * dy - sp->y;
      px -= lx * nearbyint(px / lx);
      py -= ly * nearbyint(py / ly);
      double w = exp(-(px * px + py * py) / r2);
      size_t k = (size_t)j * nx + i;
      if (sp->is_dye) {
        dye[4 * k] += sp->r * w;
        dye[4 * k + 1] += sp->g * w;
        dye[4 * k + 2] += sp->b * w;
      } else {
        u[k] += sp->dx * w;
        v[k] += sp->dy * w;
      }
    }
}

/* ------------------------------------------------------------------ diffusion */

/* src/cpu/solver.js:56-69 */
void or_diffuse_explicit(int nx, int ny, double lx, double ly, const double *f, double *out,
                         double nu, double dt) {
  double dx = lx / nx, dy = ly / ny;
  if (nu <= 0) { memcpy(out, f, sizeof(double) * nx * ny); return; }
  double alpha = nu * dt / (dx * dx), beta = nu * dt / (dy * dy);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    int jm = or_wrap(j - 1, ny) * nx, jp = or_wrap(j + 1, ny) * nx, j0 = j * nx;
    for (int i = 0; i < nx; i++) {
      int im = or_wrap(i - 1, nx), ip = or_wrap(i + 1, nx);
      out[j0 + i] = f[j0 + i] + alpha * (f[j0 + im] - 2 * f[j0 + i] + f[j0 + ip]) +
                    beta * (f[jm + i] - 2 * f[j0 + i] + f[jp + i]);
    }
  }
}

/* src/cpu/solver.js:71-111 -- CG on (I - nu dt Lap), warm-started from f */
int or_diffuse_implicit(int nx, int ny, double lx, double ly, const double *f, double *out,
                        double nu, double dt, double tol) {
  size_t n = (size_t)nx * ny;
  if (nu <= 0) { memcpy(out, f, sizeof(double) * n); return 0; }
  double k = nu * dt;
  double *r = (double *)malloc(sizeof(double) * n), *d = (double *)malloc(sizeof(double) * n);
  double *Ap = (double *)malloc(sizeof(double) * n);
  memcpy(out, f, sizeof(double) * n);
  or_laplacian(nx, ny, lx, ly, out, Ap);
  for (size_t i = 0; i < n; i++) { Ap[i] = out[i] - k * Ap[i]; r[i] = f[i] - Ap[i]; d[i] = r[i]; }
  double rr = 0, bb = 0;
  for (size_t i = 0; i < n; i++) { rr += r[i] * r[i]; bb += f[i] * f[i]; }
  double bnorm = bb > 0 ? sqrt(bb) : 1.0;
  int it = 0;
  for (; it < 400; it++) {
    if (sqrt(rr) / bnorm <= tol) break;
    or_laplacian(nx, ny, lx, ly, d, Ap);
    double dAd = 0;
    for (size_t i = 0; i < n; i++) { Ap[i] = d[i] - k * Ap[i]; dAd += d[i] * Ap[i]; }
    if (fabs(dAd) < 1e-300) break;
    double alpha = rr / dAd, rrn = 0;
    for (size_t i = 0; i < n; i++) { out[i] += alpha * d[i]; r[i] -= alpha * Ap[i]; rrn += r[i] * r[i]; }
    double beta = rrn / rr;
    for (size_t i = 0; i < n; i++) d[i] = r[i] + beta * d[i];
    rr = rrn;
  }
  free(r); free(d); free(Ap);
  return it;
}

/* ------------------------------------------------------------------------ FFT */

static int is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

/* src/core/fft.js:33-69 -- radix-2, sign -1 forward, +1 inverse (unnormalised).
 * Twiddles are evaluated directly rather than by recurrence. */
int or_fft(int n, double *re, double *im, int sign) {
  if (!is_pow2(n)) return -1;
  for (int i = 1, j = 0; i < n; i++) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) {
      double t = re[i]; re[i] = re[j]; re[j] = t;
      t = im[i]; im[i] = im[j]; im[j] = t;
    }
  }
  for (int len = 2; len <= n; len <<= 1) {
    int half = len / 2;
    for (int k = 0; k < half; k++) {
      double ang = sign * 2.0 * M_PI * k / len;
      double cr = cos(ang), ci = sin(ang);
      for (int i = 0; i < n; i += len) {
        double ar = re[i + k], ai = im[i + k];
        double br = re[i + k + half], bi = im[i + k + half];
        double dr = br * cr - bi * ci, di = br * ci + bi * cr;
        re[i + k] = ar + dr; im[i + k] = ai + di;
        re[i + k + half] = ar - dr; im[i + k + half] = ai - di;
      }
    }
  }
  return 0;
}

/* src/core/fft.js:112-160 -- rows then columns; inverse scaled by 1/(nx*ny) */
int or_fft2(int nx, int ny, double *re, double *im, int sign) {
  if (!is_pow2(nx) || !is_pow2(ny)) return -1;
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) or_fft(nx, re + (size_t)j * nx, im + (size_t)j * nx, sign);
#pragma omp parallel
  {
    double *cr = (double *)malloc(sizeof(double) * ny), *ci = (double *)malloc(sizeof(double) * ny);
#pragma omp for schedule(static)
    for (int i = 0; i < nx; i++) {
      for (int j = 0; j < ny; j++) { cr[j] = re[(size_t)j * nx + i]; ci[j] = im[(size_t)j * nx + i]; }
      or_fft(ny, cr, ci, sign);
      for (int j = 0; j < ny; j++) { re[(size_t)j * nx + i] = cr[j]; im[(size_t)j * nx + i] = ci[j]; }
    }
    free(cr); free(ci);
  }
  if (sign > 0) {
    double N = (double)nx * ny;
    for (size_t k = 0; k < (size_t)nx * ny; k++) { re[k] /= N; im[k] /= N; }
  }
  return 0;
}

/* src/cpu/projection.js:208-236 -- divide by the discrete 5-point eigenvalue */
int or_poisson_fft(int nx, int ny, double lx, double ly, const double *rhs, double *p) {
  if (!is_pow2(nx) || !is_pow2(ny)) return -1;
  double dx = lx / nx, dy = ly / ny;
  size_t n = (size_t)nx * ny;
  double *re = (double *)malloc(sizeof(double) * n), *im = (double *)calloc(n, sizeof(double));
  memcpy(re, rhs, sizeof(double) * n);
  or_fft2(nx, ny, re, im, -1);
  for (int j = 0; j < ny; j++) {
    double sy = sin(M_PI * j / ny);
    double lyv = -4 * sy * sy / (dy * dy);
    for (int i = 0; i < nx; i++) {
      double sx = sin(M_PI * i / nx);
      double lam = -4 * sx * sx / (dx * dx) + lyv;
      size_t k = (size_t)j * nx + i;
      if (i == 0 && j == 0) { re[k] = 0; im[k] = 0; }
      else { re[k] /= lam; im[k] /= lam; }
    }
  }
  or_fft2(nx, ny, re, im, +1);
  memcpy(p, re, sizeof(double) * n);
  free(re); free(im);
  return 0;
}

/* The periodic 5-point problem L x = rhs - mean(rhs) on an n x n grid of spacing
 * lx / n, solved exactly for any n: a separable complex DFT with directly evaluated
 * twiddles, division by the eigenvalue -(4/h^2)(sin^2(pi kx/n) + sin^2(pi ky/n)), the
 * inverse DFT. The solution has zero mean. O(n^3); the exact coarsest-level solve of
 * the multigrid variant that stops at a level <= 64 (not in the reference). */
static void dft_lines(int n, double *re, double *im, int stride, int step, int sign, const double *cs,
                      const double *sn, double *tr, double *ti) {
  for (int line = 0; line < n; line++) {
    double *r = re + (size_t)line * step, *q = im + (size_t)line * step;
    for (int k = 0; k < n; k++) {
      double ar = 0, ai = 0;
      for (int i = 0; i < n; i++) {
        int m = (int)(((long)i * k) % n);
        double c = cs[m], s_ = sign * sn[m];
        double xr = r[(size_t)i * stride], xi = q[(size_t)i * stride];
        ar += xr * c - xi * s_;
        ai += xr * s_ + xi * c;
      }
      tr[k] = ar;
      ti[k] = ai;
    }
    for (int k = 0; k < n; k++) { r[(size_t)k * stride] = tr[k]; q[(size_t)k * stride] = ti[k]; }
  }
}

int or_poisson_dft(int n, double lx, const double *rhs, double *x) {
  if (n < 1) return -1;
  size_t N = (size_t)n * n;
  double h = lx / n;
  double *re = (double *)malloc(sizeof(double) * N), *im = (double *)calloc(N, sizeof(double));
  double *cs = (double *)malloc(sizeof(double) * n), *sn = (double *)malloc(sizeof(double) * n);
  double *tr = (double *)malloc(sizeof(double) * n), *ti = (double *)malloc(sizeof(double) * n);
  for (int m = 0; m < n; m++) { cs[m] = cos(2 * M_PI * m / n); sn[m] = sin(2 * M_PI * m / n); }
  memcpy(re, rhs, sizeof(double) * N);
  dft_lines(n, re, im, 1, n, -1, cs, sn, tr, ti);  /* along x, one row per line */
  dft_lines(n, re, im, n, 1, -1, cs, sn, tr, ti);  /* along y, one column per line */
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      double sx = sin(M_PI * i / n), sy = sin(M_PI * j / n);
      double lam = -4.0 * (sx * sx + sy * sy) / (h * h);
      size_t k = (size_t)j * n + i;
      if (i == 0 && j == 0) { re[k] = 0; im[k] = 0; }
      else { re[k] /= lam; im[k] /= lam; }
    }
  dft_lines(n, re, im, 1, n, +1, cs, sn, tr, ti);
  dft_lines(n, re, im, n, 1, +1, cs, sn, tr, ti);
  for (size_t k = 0; k < N; k++) x[k] = re[k] / (double)N;
  free(re); free(im); free(cs); free(sn); free(tr); free(ti);
  return 0;
}

static double dotv(const double *a, const double *b, size_t n) {
  double s = 0;
  for (size_t i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}
static void remove_mean(double *a, size_t n) {
  double s = 0;
  for (size_t i = 0; i < n; i++) s += a[i];
  double m = s / n;
  for (size_t i = 0; i < n; i++) a[i] -= m;
}

/* src/cpu/projection.js:113-149 */
int or_poisson_cg(int nx, int ny, double lx, double ly, const double *rhs, double *p,
                  double tol, int maxit, double *residual_out) {
  size_t n = (size_t)nx * ny;
  double *work = (double *)malloc(sizeof(double) * n), *r = (double *)malloc(sizeof(double) * n);
  double *d = (double *)malloc(sizeof(double) * n), *Ap = (double *)malloc(sizeof(double) * n);
  memcpy(work, rhs, sizeof(double) * n);
  remove_mean(work, n);
  memcpy(r, work, sizeof(double) * n);
  or_laplacian(nx, ny, lx, ly, p, Ap);
  for (size_t i = 0; i < n; i++) r[i] -= Ap[i];
  remove_mean(r, n);
  memcpy(d, r, sizeof(double) * n);
  double rr = dotv(r, r, n);
  double rn = sqrt(dotv(work, work, n));
  double rhs_norm = rn > 0 ? rn : 1.0;
  double residual = sqrt(rr) / rhs_norm;
  int it = 0;
  if (residual > tol) {
    for (; it < maxit; it++) {
      or_laplacian(nx, ny, lx, ly, d, Ap);
      remove_mean(Ap, n);
      double dAd = dotv(d, Ap, n);
      if (fabs(dAd) < 1e-300) break;
      double alpha = rr / dAd;
      for (size_t i = 0; i < n; i++) { p[i] += alpha * d[i]; r[i] -= alpha * Ap[i]; }
      double rrn = dotv(r, r, n);
      residual = sqrt(rrn) / rhs_norm;
      if (residual <= tol) { it++; break; }
      double beta = rrn / rr;
      for (size_t i = 0; i < n; i++) d[i] = r[i] + beta * d[i];
      rr = rrn;
    }
    remove_mean(p, n);
  }
  if (residual_out) *residual_out = residual;
  free(work); free(r); free(d); free(Ap);
  return it;
}

/* ------------------------------------------- GPU pressure-solver semantics */

/* src/gpu/shaders.js:302-325 (JACOBI) with the ping-pong of src/gpu/solver.js:726-737:
 * p' = p + omega*((l+r+b+t - dx^2*rhs)/4 - p); square cells assumed (h^2 = dx^2). */
void or_jacobi(int nx, int ny, double lx, double ly, const double *rhs, double *p, int iters, double omega) {
  double dx = lx / nx, h2 = dx * dx;
  (void)ly;
  size_t n = (size_t)nx * ny;
  double *a = p, *b = (double *)malloc(sizeof(double) * n);
  for (int it = 0; it < iters; it++) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++) {
      int jm = or_wrap(j - 1, ny) * nx, jp = or_wrap(j + 1, ny) * nx, j0 = j * nx;
      for (int i = 0; i < nx; i++) {
        int im = or_wrap(i - 1, nx), ip = or_wrap(i + 1, nx);
        double jac = (a[j0 + im] + a[j0 + ip] + a[jm + i] + a[jp + i] - h2 * rhs[j0 + i]) * 0.25;
        b[j0 + i] = a[j0 + i] + omega * (jac - a[j0 + i]);
      }
    }
    double *t = a; a = b; b = t;
  }
  if (a != p) { memcpy(p, a, sizeof(double) * n); b = a; }
  free(b);
}

/* one colour pass, src/gpu/shaders.js:340-372: colour = (i+j)%2; the other
 * colour is unchanged, so an in-place update equals the reference's copy-through. */
static void rb_pass(int nx, int ny, double h2, const double *rhs, double *p, int colour, double omega) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    int jm = or_wrap(j - 1, ny) * nx, jp = or_wrap(j + 1, ny) * nx, j0 = j * nx;
    for (int i = ((j + colour) & 1); i < nx; i += 2) {
      int im = or_wrap(i - 1, nx), ip = or_wrap(i + 1, nx);
      double gs = (p[j0 + im] + p[j0 + ip] + p[jm + i] + p[jp + i] - h2 * rhs[j0 + i]) * 0.25;
      p[j0 + i] = p[j0 + i] + omega * (gs - p[j0 + i]);
    }
  }
}

/* src/gpu/solver.js:706-724 -- colour 0 then colour 1 per sweep */
void or_redblack(int nx, int ny, double lx, double ly, const double *rhs, double *p, int sweeps, double omega) {
  double dx = lx / nx;
  (void)ly;
  for (int s = 0; s < sweeps; s++) {
    rb_pass(nx, ny, dx * dx, rhs, p, 0, omega);
    rb_pass(nx, ny, dx * dx, rhs, p, 1, omega);
  }
}

/* src/gpu/multigrid.js:182-197 -- halve while n >= min_size; every level stays even
 * (an odd periodic level has same-colour neighbours at its seam) */
int or_mg_levels(int n, int min_size) {
  int L = 0;
  while (n >= min_size) {
    L++;
    if (n / 2 < min_size || (n & 1) || ((n / 2) & 1)) break;
    n /= 2;
  }
  return L;
}

typedef struct { int n; double dx; double *x, *rhs, *res; } MGL;

/* src/gpu/multigrid.js:48-66 */
static void mg_residual(const MGL *l) {
  int n = l->n;
  double idx2 = 1.0 / (l->dx * l->dx);
  for (int j = 0; j < n; j++) {
    int jm = or_wrap(j - 1, n) * n, jp = or_wrap(j + 1, n) * n, j0 = j * n;
    for (int i = 0; i < n; i++) {
      int im = or_wrap(i - 1, n), ip = or_wrap(i + 1, n);
      double lap = (l->x[j0 + im] + l->x[j0 + ip] + l->x[jm + i] + l->x[jp + i] - 4.0 * l->x[j0 + i]) * idx2;
      l->res[j0 + i] = l->rhs[j0 + i] - lap;
    }
  }
}

/* src/gpu/multigrid.js:72-93 -- 2x2 average */
static void mg_restrict(const MGL *f, MGL *c) {
  int n = c->n, fn = f->n;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      double s = f->res[(2 * j) * fn + 2 * i] + f->res[(2 * j) * fn + 2 * i + 1] +
                 f->res[(2 * j + 1) * fn + 2 * i] + f->res[(2 * j + 1) * fn + 2 * i + 1];
      c->rhs[j * n + i] = s * 0.25;
    }
}

static inline double mixd(double a, double b, double t) { return a * (1.0 - t) + b * t; }

/* src/gpu/multigrid.js:99-131 -- bilinear cell-centred, added to the fine field */
static void mg_prolong(const MGL *c, MGL *f) {
  int n = f->n, cn = c->n;
  for (int j = 0; j < n; j++) {
    double gy = (j + 0.5) * 0.5 - 0.5;
    int j0 = (int)floor(gy);
    double fy = gy - floor(gy);
    int y0 = or_wrap(j0, cn), y1 = or_wrap(j0 + 1, cn);
    for (int i = 0; i < n; i++) {
      double gx = (i + 0.5) * 0.5 - 0.5;
      int i0 = (int)floor(gx);
      double fx = gx - floor(gx);
      int x0 = or_wrap(i0, cn), x1 = or_wrap(i0 + 1, cn);
      double a = c->x[y0 * cn + x0], b = c->x[y0 * cn + x1];
      double cc = c->x[y1 * cn + x0], d = c->x[y1 * cn + x1];
      double corr = mixd(mixd(a, b, fx), mixd(cc, d, fx), fy);
      f->x[j * n + i] += corr;
    }
  }
}

/* src/gpu/multigrid.js:256-270 */
static void mg_smooth(MGL *l, int sweeps) {
  double h2 = l->dx * l->dx;
  for (int s = 0; s < sweeps; s++) {
    rb_pass(l->n, l->n, h2, l->rhs, l->x, 0, 1.0);
    rb_pass(l->n, l->n, h2, l->rhs, l->x, 1, 1.0);
  }
}

/* src/gpu/multigrid.js:277-330 (V-cycle) and src/gpu/solver.js:686-703 (k cycles
 * accumulating into the finest solution). p is the finest x on entry and exit;
 * history[k] (if given) receives ||Lp - rhs|| / ||rhs|| after cycle k+1. */
int or_mg_vcycles(int n, double lx, const double *rhs, double *p, int cycles,
                  int min_size, int pre, int post, int coarse, double *history) {
  int L = or_mg_levels(n, min_size);
  if (L < 1) return -1;
  MGL *lv = (MGL *)calloc(L, sizeof(MGL));
  int m = n;
  for (int i = 0; i < L; i++) {
    size_t c = (size_t)m * m;
    lv[i].n = m;
    lv[i].dx = lx / m;
    lv[i].x = i == 0 ? p : (double *)calloc(c, sizeof(double));
    lv[i].rhs = i == 0 ? (double *)rhs : (double *)calloc(c, sizeof(double));
    lv[i].res = (double *)calloc(c, sizeof(double));
    m /= 2;
  }
  for (int k = 0; k < cycles; k++) {
    for (int i = 0; i < L - 1; i++) {
      mg_smooth(&lv[i], pre);
      mg_residual(&lv[i]);
      mg_restrict(&lv[i], &lv[i + 1]);
      memset(lv[i + 1].x, 0, sizeof(double) * lv[i + 1].n * lv[i + 1].n);
    }
    mg_smooth(&lv[L - 1], coarse);
    for (int i = L - 2; i >= 0; i--) {
      mg_prolong(&lv[i + 1], &lv[i]);
      mg_smooth(&lv[i], post);
    }
    if (history) history[k] = or_poisson_residual_rel(n, n, lx, lx, p, rhs);
  }
  for (int i = 0; i < L; i++) {
    if (i > 0) { free(lv[i].x); free(lv[i].rhs); }
    free(lv[i].res);
  }
  free(lv);
  return L;
}

/* ---------------------------------------------------------- lid-driven cavity */
/* Unit box [0,1]^2, n x n cells, h = 1/n, the periodic MAC layout: u[j n + i] on the
 * left face of cell (i, j), v[j n + i] on its bottom face. u[j n + 0] and v[i] (row 0)
 * are wall faces and stay 0; the right and top wall faces are not stored (0). No slip
 * on every wall, the top wall (lid) moving with velocity U. Beyond a wall, ghosts put
 * the wall value on the wall: u below the bottom -u, above the lid 2U - u; v beyond a
 * side wall -v. Departure points are clamped into the box. (The reference has no
 * walls; this is the walled counterpart of the periodic operators above.) */

static inline double cav_u(const double *u, int n, double U, int i, int j) {
  if (i <= 0 || i >= n) return 0.0;
  if (j < 0) return -u[i];
  if (j >= n) return 2.0 * U - u[(size_t)(n - 1) * n + i];
  return u[(size_t)j * n + i];
}
static inline double cav_v(const double *v, int n, int i, int j) {
  if (j <= 0 || j >= n) return 0.0;
  if (i < 0) return -v[(size_t)j * n];
  if (i >= n) return -v[(size_t)j * n + n - 1];
  return v[(size_t)j * n + i];
}

/* bilinear sample of u (or v) at (x, y); lo/hi receive the four source values' range */
static double cav_sample(const double *a, int n, double U, int comp, double x, double y, double *lo, double *hi) {
  double h = 1.0 / n;
  x = x < 0 ? 0 : (x > 1 ? 1 : x);
  y = y < 0 ? 0 : (y > 1 ? 1 : y);
  double gx = comp == 0 ? x / h : x / h - 0.5;
  double gy = comp == 0 ? y / h - 0.5 : y / h;
  int i0 = (int)floor(gx), j0 = (int)floor(gy);
  if (i0 > n - 1) i0 = n - 1;
  if (j0 > n - 1) j0 = n - 1;
  double fx = gx - i0, fy = gy - j0;
  double s00, s10, s01, s11;
  if (comp == 0) {
    s00 = cav_u(a, n, U, i0, j0); s10 = cav_u(a, n, U, i0 + 1, j0);
    s01 = cav_u(a, n, U, i0, j0 + 1); s11 = cav_u(a, n, U, i0 + 1, j0 + 1);
  } else {
    s00 = cav_v(a, n, i0, j0); s10 = cav_v(a, n, i0 + 1, j0);
    s01 = cav_v(a, n, i0, j0 + 1); s11 = cav_v(a, n, i0 + 1, j0 + 1);
  }
  if (lo) {
    *lo = fmin(fmin(s00, s10), fmin(s01, s11));
    *hi = fmax(fmax(s00, s10), fmax(s01, s11));
  }
  double bot = s00 + (s10 - s00) * fx, top = s01 + (s11 - s01) * fx;
  return bot + (top - bot) * fy;
}

double or_cav_sample_u(int n, double U, const double *u, double x, double y) {
  return cav_sample(u, n, U, 0, x, y, NULL, NULL);
}
double or_cav_sample_v(int n, const double *v, double x, double y) {
  return cav_sample(v, n, 0, 1, x, y, NULL, NULL);
}

/* RK2 midpoint back-trace through (u, v) (src/cpu/advect.js:46-54 with walls) */
static void cav_trace(const double *u, const double *v, int n, double U, double x, double y, double dt,
                      double *xo, double *yo) {
  double u1 = cav_sample(u, n, U, 0, x, y, NULL, NULL), v1 = cav_sample(v, n, 0, 1, x, y, NULL, NULL);
  double xm = x - 0.5 * dt * u1, ym = y - 0.5 * dt * v1;
  double u2 = cav_sample(u, n, U, 0, xm, ym, NULL, NULL), v2 = cav_sample(v, n, 0, 1, xm, ym, NULL, NULL);
  *xo = x - dt * u2;
  *yo = y - dt * v2;
}

/* semi-Lagrangian or limited MacCormack self-advection of the interior faces
 * (src/cpu/advect.js:74-111, :149-206 with walls); wall faces stay 0 */
void or_cav_advect(int n, double U, int scheme, const double *u, const double *v, double *ou, double *ov,
                   double dt) {
  double h = 1.0 / n;
  size_t N = (size_t)n * n;
  double *fu = (double *)calloc(N, sizeof(double)), *fv = (double *)calloc(N, sizeof(double));
  double *lou = (double *)calloc(N, sizeof(double)), *hiu = (double *)calloc(N, sizeof(double));
  double *lov = (double *)calloc(N, sizeof(double)), *hiv = (double *)calloc(N, sizeof(double));
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      size_t k = (size_t)j * n + i;
      double xd, yd;
      if (i > 0) {
        cav_trace(u, v, n, U, i * h, (j + 0.5) * h, dt, &xd, &yd);
        fu[k] = cav_sample(u, n, U, 0, xd, yd, &lou[k], &hiu[k]);
      }
      if (j > 0) {
        cav_trace(u, v, n, U, (i + 0.5) * h, j * h, dt, &xd, &yd);
        fv[k] = cav_sample(v, n, 0, 1, xd, yd, &lov[k], &hiv[k]);
      }
    }
  if (scheme == OR_ADV_SL) {
    memcpy(ou, fu, sizeof(double) * N);
    memcpy(ov, fv, sizeof(double) * N);
  } else {
    /* back-trace of the forward state through the original velocity */
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        size_t k = (size_t)j * n + i;
        double xd, yd;
        ou[k] = 0;
        ov[k] = 0;
        if (i > 0) {
          cav_trace(u, v, n, U, i * h, (j + 0.5) * h, -dt, &xd, &yd);
          double back = cav_sample(fu, n, U, 0, xd, yd, NULL, NULL);
          double c = fu[k] + 0.5 * (u[k] - back);
          ou[k] = c < lou[k] ? lou[k] : (c > hiu[k] ? hiu[k] : c);
        }
        if (j > 0) {
          cav_trace(u, v, n, U, (i + 0.5) * h, j * h, -dt, &xd, &yd);
          double back = cav_sample(fv, n, 0, 1, xd, yd, NULL, NULL);
          double c = fv[k] + 0.5 * (v[k] - back);
          ov[k] = c < lov[k] ? lov[k] : (c > hiv[k] ? hiv[k] : c);
        }
      }
  }
  free(fu); free(fv); free(lou); free(hiu); free(lov); free(hiv);
}

/* explicit diffusion of the interior faces with the wall ghosts (src/cpu/solver.js:52-69) */
void or_cav_diffuse(int n, double U, const double *u, const double *v, double *ou, double *ov, double nu,
                    double dt) {
  double h = 1.0 / n, a = nu * dt / (h * h);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      size_t k = (size_t)j * n + i;
      double uc = cav_u(u, n, U, i, j), vc = cav_v(v, n, i, j);
      ou[k] = i > 0 ? uc + a * (cav_u(u, n, U, i - 1, j) + cav_u(u, n, U, i + 1, j) + cav_u(u, n, U, i, j - 1) +
                                cav_u(u, n, U, i, j + 1) - 4.0 * uc)
                    : 0.0;
      ov[k] = j > 0 ? vc + a * (cav_v(v, n, i - 1, j) + cav_v(v, n, i + 1, j) + cav_v(v, n, i, j - 1) +
                                cav_v(v, n, i, j + 1) - 4.0 * vc)
                    : 0.0;
    }
}

void or_cav_divergence(int n, const double *u, const double *v, double *div) {
  double h = 1.0 / n;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++)
      div[(size_t)j * n + i] = (cav_u(u, n, 0, i + 1, j) - cav_u(u, n, 0, i, j)) / h +
                               (cav_v(v, n, i, j + 1) - cav_v(v, n, i, j)) / h;
}

/* Laplacian with zero normal derivative at the walls (mirror ghosts) */
void or_cav_laplacian(int n, const double *p, double *out) {
  double h2 = 1.0 / ((double)n * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      size_t k = (size_t)j * n + i;
      double s = 0;
      if (i > 0) s += p[k - 1] - p[k];
      if (i < n - 1) s += p[k + 1] - p[k];
      if (j > 0) s += p[k - n] - p[k];
      if (j < n - 1) s += p[k + n] - p[k];
      out[k] = s / h2;
    }
}

/* subtract grad p on the interior faces */
void or_cav_subtract_gradient(int n, double *u, double *v, const double *p) {
  double h = 1.0 / n;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      size_t k = (size_t)j * n + i;
      if (i > 0) u[k] -= (p[k] - p[k - 1]) / h;
      if (j > 0) v[k] -= (p[k] - p[k - n]) / h;
    }
}

/* exact Neumann Poisson solve by the cosine eigenbasis phi_k(i) = cos(pi k (i + 1/2) / n)
 * (eigenvalues -(4/h^2) sin^2(pi k / 2n) per axis); the rhs mean (k = l = 0) is dropped */
int or_cav_poisson_exact(int n, const double *rhs, double *p) {
  double h = 1.0 / n;
  double *C = (double *)malloc(sizeof(double) * n * n);   /* C[k n + i] = phi_k(i) */
  double *T = (double *)malloc(sizeof(double) * n * n), *A = (double *)malloc(sizeof(double) * n * n);
  double *lam = (double *)malloc(sizeof(double) * n);
  for (int k = 0; k < n; k++) {
    lam[k] = -4.0 / (h * h) * pow(sin(M_PI * k / (2.0 * n)), 2);
    for (int i = 0; i < n; i++) C[(size_t)k * n + i] = cos(M_PI * k * (i + 0.5) / n);
  }
  /* A[l][k] = sum_j sum_i rhs[j][i] C[k][i] C[l][j] / norms */
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int k = 0; k < n; k++) {
      double s = 0;
      for (int i = 0; i < n; i++) s += rhs[(size_t)j * n + i] * C[(size_t)k * n + i];
      T[(size_t)j * n + k] = s;
    }
#pragma omp parallel for schedule(static)
  for (int l = 0; l < n; l++)
    for (int k = 0; k < n; k++) {
      double s = 0;
      for (int j = 0; j < n; j++) s += T[(size_t)j * n + k] * C[(size_t)l * n + j];
      double nk = k == 0 ? n : n / 2.0, nl = l == 0 ? n : n / 2.0;
      double den = lam[k] + lam[l];
      A[(size_t)l * n + k] = (k == 0 && l == 0) ? 0.0 : s / (nk * nl) / den;
    }
#pragma omp parallel for schedule(static)
  for (int l = 0; l < n; l++)
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int k = 0; k < n; k++) s += A[(size_t)l * n + k] * C[(size_t)k * n + i];
      T[(size_t)l * n + i] = s;
    }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      double s = 0;
      for (int l = 0; l < n; l++) s += T[(size_t)l * n + i] * C[(size_t)l * n + j];
      p[(size_t)j * n + i] = s;
    }
  free(C); free(T); free(A); free(lam);
  return 0;
}

/* disk coupling with plain separations, wall faces untouched */
void or_cav_apply_coupling(int n, double *u, double *v, const or_disk *s, double mask_width, int passes,
                           double *imp) {
  double h = 1.0 / n, dA = h * h, width = mask_width * h;
  double fx = 0, fy = 0, tq = 0;
  for (int pass = 0; pass < passes; pass++) {
    for (int j = 0; j < n; j++)
      for (int i = 1; i < n; i++) {
        double px = i * h, py = (j + 0.5) * h, dx = px - s->x, dy = py - s->y;
        double chi = 0.5 * (1.0 - tanh((hypot(dx, dy) - s->r) / width));
        if (chi < 1e-6) continue;
        size_t k = (size_t)j * n + i;
        double du = chi * ((s->vx - s->omega * dy) - u[k]);
        u[k] += du;
        fx += du * dA;
        tq += -dy * du * dA;
      }
    for (int j = 1; j < n; j++)
      for (int i = 0; i < n; i++) {
        double px = (i + 0.5) * h, py = j * h, dx = px - s->x, dy = py - s->y;
        double chi = 0.5 * (1.0 - tanh((hypot(dx, dy) - s->r) / width));
        if (chi < 1e-6) continue;
        size_t k = (size_t)j * n + i;
        double dv = chi * ((s->vy + s->omega * dx) - v[k]);
        v[k] += dv;
        fy += dv * dA;
        tq += dx * dv * dA;
      }
  }
  if (imp) { imp[0] = fx; imp[1] = fy; imp[2] = tq; }
}

/* one cavity step: advect, diffuse, couple the disk, project exactly */
int or_cav_step(int n, double U, double nu, int scheme, const or_disk *disk, double mask_width, int passes,
                double dt, double *u, double *v, double *p_out) {
  size_t N = (size_t)n * n;
  double *tu = (double *)malloc(sizeof(double) * N), *tv = (double *)malloc(sizeof(double) * N);
  double *div = (double *)malloc(sizeof(double) * N), *p = (double *)malloc(sizeof(double) * N);
  or_cav_advect(n, U, scheme, u, v, tu, tv, dt);
  if (nu > 0) or_cav_diffuse(n, U, tu, tv, u, v, nu, dt);
  else { memcpy(u, tu, sizeof(double) * N); memcpy(v, tv, sizeof(double) * N); }
  if (disk) or_cav_apply_coupling(n, u, v, disk, mask_width, passes, NULL);
  or_cav_divergence(n, u, v, div);
  or_cav_poisson_exact(n, div, p);
  or_cav_subtract_gradient(n, u, v, p);
  if (p_out) memcpy(p_out, p, sizeof(double) * N);
  free(tu); free(tv); free(div); free(p);
  return 0;
}

/* -------------------------------------- refinement with stored-precision rounding */

/* value after a store in `storage` (0: FP32, 1: BF16) of the FP32 rounding of x;
 * BF16 keeps the upper 16 bits of the FP32, rounded to nearest even */
double or_round_storage(double x, int storage) {
  float f = (float)x;
  if (storage == 1) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
  }
  return (double)f;
}

static void round_field(double *a, size_t n, int storage) {
  for (size_t k = 0; k < n; k++) a[k] = or_round_storage(a[k], storage);
}

/* The projection as the GPU solver runs it with a fused V-cycle: iterative
 * refinement p <- p + e with r = div - L p in FP64, each V-cycle on the
 * correction e from zero, and the right-hand sides and pre-smoothed corrections
 * rounded to the storage precision wherever they pass through memory: the level-0
 * residual, and the pre-smoothed correction and restricted rhs of every level above
 * coarse_n. The corrections passed up (the coarse solve's result and each
 * post-smoothed correction) are stored in FP32 under every storage precision.
 * Levels at or below coarse_n run unrounded (they stay on chip). The V-cycle is that of
 * src/gpu/multigrid.js:277-330; with no rounding the k-th p equals k plain
 * V-cycles (the cycle is affine in its right-hand side).
 * coarse_exact: the hierarchy stops at its first level <= coarse_n, and that level is
 * solved exactly (or_poisson_dft) instead of by sweeps and the levels below it.
 * rmax[it] = max|div - L p| before refinement step it, for it = 0..iterations. */
int or_mg_refine(int n, double lx, const double *div, double *p, int iterations, int min_size, int pre,
                 int post, int coarse, int coarse_n, int storage, int coarse_exact, double *rmax) {
  int L = or_mg_levels(n, min_size);
  if (L < 1) return -1;
  if (coarse_exact) {
    /* the hierarchy stops at its first level <= coarse_n, which is solved exactly */
    int m = n, l = 1;
    while (l < L && m > coarse_n) { m /= 2; l++; }
    if (m > coarse_n) return -2;
    L = l;
  }
  MGL *lv = (MGL *)calloc(L, sizeof(MGL));
  double **xs = (double **)calloc(L, sizeof(double *));  /* stored pre-smoothed corrections */
  int m = n, cs = L;
  for (int i = 0; i < L; i++) {
    size_t c = (size_t)m * m;
    lv[i].n = m;
    lv[i].dx = lx / m;
    lv[i].x = (double *)calloc(c, sizeof(double));
    lv[i].rhs = (double *)calloc(c, sizeof(double));
    lv[i].res = (double *)calloc(c, sizeof(double));
    xs[i] = (double *)calloc(c, sizeof(double));
    if (cs == L && m <= coarse_n) cs = i;
    m /= 2;
  }
  if (cs >= L) cs = L - 1;  /* the coarsest level is always solved on chip */
  size_t N = (size_t)n * n;
  double *lap = (double *)malloc(sizeof(double) * N);
  double *e = (double *)calloc(N, sizeof(double));
  for (int it = 0; it <= iterations; it++) {
    if (it > 0)
      for (size_t k = 0; k < N; k++) p[k] += e[k];
    or_laplacian(n, n, lx, lx, p, lap);
    double mx = 0;
    for (size_t k = 0; k < N; k++) {
      double r = div[k] - lap[k];
      if (fabs(r) > mx) mx = fabs(r);
      lv[0].rhs[k] = or_round_storage(r, storage);
    }
    if (rmax) rmax[it] = mx;
    if (it == iterations) break;
    /* grid-wide levels: smooth from zero, restrict, store */
    for (int i = 0; i < cs; i++) {
      size_t c = (size_t)lv[i].n * lv[i].n;
      memset(lv[i].x, 0, sizeof(double) * c);
      mg_smooth(&lv[i], pre);
      mg_residual(&lv[i]);
      mg_restrict(&lv[i], &lv[i + 1]);
      round_field(lv[i + 1].rhs, (size_t)lv[i + 1].n * lv[i + 1].n, storage);
      memcpy(xs[i], lv[i].x, sizeof(double) * c);
      round_field(xs[i], c, storage);
    }
    /* on-chip levels cs..L-1: a plain V-cycle from zero */
    for (int i = cs; i < L; i++) memset(lv[i].x, 0, sizeof(double) * lv[i].n * lv[i].n);
    for (int i = cs; i < L - 1; i++) {
      mg_smooth(&lv[i], pre);
      mg_residual(&lv[i]);
      mg_restrict(&lv[i], &lv[i + 1]);
    }
    /* the coarsest rhs loses its mean first (the periodic compatibility condition,
     * src/cpu/projection.js:91-102): stored rounding leaves it slightly non-zero,
     * and sweeps on an incompatible rhs drift by a growing constant */
    remove_mean(lv[L - 1].rhs, (size_t)lv[L - 1].n * lv[L - 1].n);
    if (coarse_exact) or_poisson_dft(lv[L - 1].n, lv[L - 1].n * lv[L - 1].dx, lv[L - 1].rhs, lv[L - 1].x);
    else mg_smooth(&lv[L - 1], coarse);
    for (int i = L - 2; i >= cs; i--) {
      mg_prolong(&lv[i + 1], &lv[i]);
      mg_smooth(&lv[i], post);
    }
    /* grid-wide levels up: stored x + prolongated correction, post-smooth, store */
    for (int i = cs - 1; i >= 0; i--) {
      size_t c = (size_t)lv[i].n * lv[i].n;
      memcpy(lv[i].x, xs[i], sizeof(double) * c);
      mg_prolong(&lv[i + 1], &lv[i]);
      mg_smooth(&lv[i], post);
    }
    memcpy(e, lv[0].x, sizeof(double) * N);
  }
  for (int i = 0; i < L; i++) { free(lv[i].x); free(lv[i].rhs); free(lv[i].res); free(xs[i]); }
  free(lv); free(xs); free(lap); free(e);
  return L;
}

/* -------------------------------------------------------------------- forcing */

static double js_round(double x) {
  double r = floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}

/* src/cpu/solver.js:179-199 -- 24 modes from the exact uint32 LCG */
int or_band_modes(uint32_t seed, double kf, double width, int *kx, int *ky, double *phase) {
  uint32_t s = seed;
#define RND() (s = s * 1664525u + 1013904223u, (double)s / 4294967296.0)
  for (int m = 0; m < 24; m++) {
    double theta = 2 * M_PI * RND();
    double kmag = kf + (RND() - 0.5) * 2 * width;
    kx[m] = (int)js_round(kmag * cos(theta));
    ky[m] = (int)js_round(kmag * sin(theta));
    phase[m] = 2 * M_PI * RND();
  }
#undef RND
  return 24;
}

/* src/cpu/solver.js:200-222 -- curl of a streamfunction, added at the faces */
void or_band_forcing(int nx, int ny, double lx, double ly, double *u, double *v,
                     double kf, double width, double amp, uint32_t seed, double dt) {
  int kx[24], ky[24];
  double ph[24];
  int M = or_band_modes(seed, kf, width, kx, ky, ph);
  double dx = lx / nx, dy = ly / ny;
  double scale = amp * dt / sqrt((double)M);
  for (int m = 0; m < M; m++) {
    if (kx[m] == 0 && ky[m] == 0) continue;
    double kk = (double)kx[m] * kx[m] + (double)ky[m] * ky[m];
    double cu = scale * (ky[m] / sqrt(kk)), cv = scale * (-kx[m] / sqrt(kk));
#pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; j++) {
      for (int i = 0; i < nx; i++) {
        double xu = i * dx, yu = (j + 0.5) * dy;
        u[j * nx + i] += cu * cos(kx[m] * xu + ky[m] * yu + ph[m]);
        double xv = (i + 0.5) * dx, yv = j * dy;
        v[j * nx + i] += cv * cos(kx[m] * xv + ky[m] * yv + ph[m]);
      }
    }
  }
}

/* FORCING, src/gpu/shaders.js:532-566, with the parameters _encodeForcing writes
 * (src/gpu/solver.js:387-396): drag on the old velocity, then the force */
void or_mode_forcing(int nx, int ny, double lx, double ly, double *u, double *v, int count, const double *kx,
                     const double *ky, const double *phase, double amp, double drag, double drift, double t,
                     double dt) {
  double dx = lx / nx, dy = ly / ny;
  double scale = amp * dt / sqrt((double)count);
  double damp = exp(-drag * dt);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++) {
    for (int i = 0; i < nx; i++) {
      double xu = i * dx, yu = (j + 0.5) * dy, xv = (i + 0.5) * dx, yv = j * dy;
      double fu = 0, fv = 0;
      for (int m = 0; m < count; m++) {
        double kk = sqrt(kx[m] * kx[m] + ky[m] * ky[m]);
        if (kk < 1e-6) continue;
        double ph = phase[m] + drift * t;
        fu += (ky[m] / kk) * cos(kx[m] * xu + ky[m] * yu + ph);
        fv -= (kx[m] / kk) * cos(kx[m] * xv + ky[m] * yv + ph);
      }
      u[j * nx + i] = u[j * nx + i] * damp + scale * fu;
      v[j * nx + i] = v[j * nx + i] * damp + scale * fv;
    }
  }
}

/* src/cpu/solver.js:233-237 */
void or_linear_drag(int n, double *u, double *v, double alpha, double dt) {
  if (alpha <= 0) return;
  double f = exp(-alpha * dt);
  for (int i = 0; i < n; i++) { u[i] *= f; v[i] *= f; }
}

/* ---------------------------------------------------------------- closed forms */

/* src/core/analytic.js:38-72 */
void or_taylor_green(int nx, int ny, double lx, double ly, double t, double nu, double *u, double *v) {
  double dx = lx / nx, dy = ly / ny, a = exp(-2 * nu * t);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      double xu = i * dx, yu = (j + 0.5) * dy, xv = (i + 0.5) * dx, yv = j * dy;
      u[j * nx + i] = sin(xu) * cos(yu) * a;
      v[j * nx + i] = -cos(xv) * sin(yv) * a;
    }
}

/* src/core/analytic.js:195-206 */
double or_velocity_error_tg(int nx, int ny, double lx, double ly, const double *u, const double *v,
                            double t, double nu) {
  size_t n = (size_t)nx * ny;
  double *eu = (double *)malloc(sizeof(double) * n), *ev = (double *)malloc(sizeof(double) * n);
  or_taylor_green(nx, ny, lx, ly, t, nu, eu, ev);
  double num = 0, den = 0;
  for (size_t k = 0; k < n; k++) {
    double du = u[k] - eu[k], dv = v[k] - ev[k];
    num += du * du + dv * dv;
    den += eu[k] * eu[k] + ev[k] * ev[k];
  }
  free(eu); free(ev);
  if (den < 1e-300) return sqrt(num / (2.0 * n));
  return sqrt(num / den);
}

/* src/core/fft.js:196-234 */
int or_energy_spectrum(int nx, int ny, const double *cu, const double *cv, double *E, int *counts) {
  if (!is_pow2(nx) || !is_pow2(ny)) return -1;
  size_t n = (size_t)nx * ny;
  double N = (double)n;
  double *ur = (double *)malloc(sizeof(double) * n), *ui = (double *)calloc(n, sizeof(double));
  double *vr = (double *)malloc(sizeof(double) * n), *vi = (double *)calloc(n, sizeof(double));
  memcpy(ur, cu, sizeof(double) * n);
  memcpy(vr, cv, sizeof(double) * n);
  or_fft2(nx, ny, ur, ui, -1);
  or_fft2(nx, ny, vr, vi, -1);
  int kmax = (nx < ny ? nx : ny) / 2;
  for (int k = 0; k <= kmax; k++) { E[k] = 0; counts[k] = 0; }
  for (int j = 0; j < ny; j++) {
    int ky = j <= ny / 2 ? j : j - ny;
    for (int i = 0; i < nx; i++) {
      int kx = i <= nx / 2 ? i : i - nx;
      size_t id = (size_t)j * nx + i;
      double pu = (ur[id] * ur[id] + ui[id] * ui[id]) / (N * N);
      double pv = (vr[id] * vr[id] + vi[id] * vi[id]) / (N * N);
      double e = 0.5 * (pu + pv);
      int shell = (int)js_round(sqrt((double)kx * kx + (double)ky * ky));
      if (shell <= kmax) { E[shell] += e; counts[shell]++; }
    }
  }
  free(ur); free(ui); free(vr); free(vi);
  return kmax;
}

/* ---------------------------------------------------------- immersed disk */

/* shortest signed separation on a periodic axis (src/cpu/solid.js:38-40);
 * Math.round rounds halves up, so floor(x + 0.5) */
double or_wrap_delta(double d, double l) { return d - l * floor(d / l + 0.5); }

/* smoothed indicator 0.5 (1 - tanh((|x - c| - r) / width)) (src/cpu/solid.js:101-107) */
double or_disk_mask(const or_disk *s, double px, double py, double lx, double ly, double width) {
  double dx = or_wrap_delta(px - s->x, lx), dy = or_wrap_delta(py - s->y, ly);
  double d = hypot(dx, dy) - s->r;
  return 0.5 * (1.0 - tanh(d / width));
}

/* rigid-body velocity: translation + omega x r (src/cpu/solid.js:110-115) */
void or_disk_velocity(const or_disk *s, double px, double py, double lx, double ly, double *vx, double *vy) {
  double dx = or_wrap_delta(px - s->x, lx), dy = or_wrap_delta(py - s->y, ly);
  *vx = s->vx - s->omega * dy;
  *vy = s->vy + s->omega * dx;
}

/* direct forcing u += chi (U_s - u) on u-faces then v-faces, `passes` times,
 * faces with chi < 1e-6 skipped (src/cpu/solid.js:165-233) */
void or_apply_coupling(int nx, int ny, double lx, double ly, double *u, double *v, const or_disk *s,
                       double mask_width, int passes, double *imp) {
  double hx = lx / nx, hy = ly / ny, dA = hx * hy, width = mask_width * hx;
  double fx = 0, fy = 0, tq = 0;
  for (int pass = 0; pass < passes; pass++) {
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double px = i * hx, py = (j + 0.5) * hy;
        double chi = or_disk_mask(s, px, py, lx, ly, width);
        if (chi < 1e-6) continue;
        double su, sv;
        or_disk_velocity(s, px, py, lx, ly, &su, &sv);
        size_t k = (size_t)j * nx + i;
        double du = chi * (su - u[k]);
        u[k] += du;
        fx += du * dA;
        tq += -or_wrap_delta(py - s->y, ly) * du * dA;
      }
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double px = (i + 0.5) * hx, py = j * hy;
        double chi = or_disk_mask(s, px, py, lx, ly, width);
        if (chi < 1e-6) continue;
        double su, sv;
        or_disk_velocity(s, px, py, lx, ly, &su, &sv);
        size_t k = (size_t)j * nx + i;
        double dv = chi * (sv - v[k]);
        v[k] += dv;
        fy += dv * dA;
        tq += or_wrap_delta(px - s->x, lx) * dv * dA;
      }
  }
  if (imp) { imp[0] = fx; imp[1] = fy; imp[2] = tq; }
}

/* ------------------------------------------------------------------- the step */

/* src/cpu/solver.js:336-398: advect -> diffuse -> force -> drag -> solid -> project */
int or_step(const or_step_cfg *c, double *u, double *v, double *p_out, double dt, uint32_t force_seed) {
  int nx = c->nx, ny = c->ny;
  size_t n = (size_t)nx * ny;
  double *tu = (double *)malloc(sizeof(double) * n), *tv = (double *)malloc(sizeof(double) * n);
  double *div = (double *)malloc(sizeof(double) * n), *p = (double *)calloc(n, sizeof(double));
  int rc = 0;
  if (c->advection == OR_ADV_MC) or_advect_velocity_mc(nx, ny, c->lx, c->ly, u, v, tu, tv, dt);
  else if (c->advection == OR_ADV_SL) or_advect_velocity_sl(nx, ny, c->lx, c->ly, u, v, tu, tv, dt);
  else { rc = -1; goto done; }
  if (c->nu > 0) {
    if (c->explicit_diffusion) {
      or_diffuse_explicit(nx, ny, c->lx, c->ly, tu, u, c->nu, dt);
      or_diffuse_explicit(nx, ny, c->lx, c->ly, tv, v, c->nu, dt);
    } else {
      or_diffuse_implicit(nx, ny, c->lx, c->ly, tu, u, c->nu, dt, 1e-12);
      or_diffuse_implicit(nx, ny, c->lx, c->ly, tv, v, c->nu, dt, 1e-12);
    }
  } else {
    memcpy(u, tu, sizeof(double) * n);
    memcpy(v, tv, sizeof(double) * n);
  }
  if (c->modes > 0) {
    or_mode_forcing(nx, ny, c->lx, c->ly, u, v, c->modes, c->mkx, c->mky, c->mph, c->mamp, c->mdrag, c->mdrift,
                    c->mt, dt);
  } else {
    if (c->forcing)
      or_band_forcing(nx, ny, c->lx, c->ly, u, v, c->kf, c->width, c->amp, force_seed, dt);
    if (c->drag > 0) or_linear_drag((int)n, u, v, c->drag, dt);
  }
  if (c->solid) {
    or_disk d = {c->sx, c->sy, c->sr, c->svx, c->svy, c->somega};
    or_apply_coupling(nx, ny, c->lx, c->ly, u, v, &d, c->mask_width, c->passes, NULL);
  }

  if (c->projection != OR_PROJ_NONE) {
    or_divergence(nx, ny, c->lx, c->ly, u, v, div);
    switch (c->projection) {
      case OR_PROJ_FFT: rc = or_poisson_fft(nx, ny, c->lx, c->ly, div, p); break;
      case OR_PROJ_CG: or_poisson_cg(nx, ny, c->lx, c->ly, div, p, c->cg_tol, c->cg_maxit, NULL); break;
      case OR_PROJ_JACOBI: or_jacobi(nx, ny, c->lx, c->ly, div, p, c->iterations, 1.0); break;
      case OR_PROJ_REDBLACK: or_redblack(nx, ny, c->lx, c->ly, div, p, c->iterations, 1.0); break;
      case OR_PROJ_MULTIGRID:
        if (nx != ny) { rc = -2; goto done; }
        if (or_mg_vcycles(nx, c->lx, div, p, c->iterations, c->mg_min_size, c->mg_pre, c->mg_post,
                          c->mg_coarse, NULL) < 1) { rc = -2; goto done; }
        break;
      default: rc = -1; goto done;
    }
    if (rc) goto done;
    or_subtract_gradient(nx, ny, c->lx, c->ly, u, v, p, 1.0);
  }
  if (p_out) memcpy(p_out, p, sizeof(double) * n);
done:
  free(tu); free(tv); free(div); free(p);
  return rc;
}
