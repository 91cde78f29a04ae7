// fs_solver.hip -- host runtime and C ABI of the MI355X stable-fluids solver.
//
// One solver owns one HIP stream (default priority) and all of its device
// memory; fs_step only enqueues. The step follows src/gpu/solver.js:518-591
// (advect -> diffuse -> force/drag -> project) with the CPU forcing semantics
// of src/cpu/solver.js:353-358 (force, then drag).
//
// Projection modes
//   jacobi, red-black   FP32 pressure, fixed sweep count, p zeroed per step
//                       (src/gpu/solver.js:658-738).
//   multigrid           iterative refinement: r = div u* - L p in FP64, one
//                       V-cycle on the FP32 correction e (L e = r), p += e in
//                       FP64. A V-cycle is affine in its start value, so k
//                       refinement iterations equal the reference's k V-cycles
//                       on p (src/gpu/solver.js:686-703) up to rounding; with a
//                       tolerance the loop stops once max|r| <= tol.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/fluidsim.h"
#include "fs_kernels.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                              \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess)                                                                         \
      return fail(e_ == hipErrorOutOfMemory ? FS_E_OOM : FS_E_DEVICE, "%s failed: %s (%s:%d)", #expr, \
                  hipGetErrorString(e_), __FILE__, __LINE__);                                    \
  } while (0)

#define LAUNCHCHK()                                                                              \
  do {                                                                                           \
    hipError_t e_ = hipGetLastError();                                                           \
    if (e_ != hipSuccess)                                                                        \
      return fail(FS_E_DEVICE, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                                     \
  } while (0)

struct Level {
  int n = 0;
  double dx = 0;
  float* x = nullptr;    // correction (level 0: e)
  float* rhs = nullptr;  // right-hand side (level 0: r)
};

struct Profile {
  int kind = -1;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<hipEvent_t> pool;
  double total_ms = 0;
  int64_t launches = 0;
};

}  // namespace

struct fs_solver_t {
  fs_config cfg{};
  int dev = 0;
  hipStream_t stream = nullptr;
  int nx = 0, ny = 0;
  size_t n = 0;
  double dx = 0, dy = 0;
  float* u[2] = {nullptr, nullptr};
  float* v[2] = {nullptr, nullptr};
  int cur = 0;
  float* fu = nullptr;
  float* fv = nullptr;
  float* div = nullptr;
  float* p32[2] = {nullptr, nullptr};
  double* p64[2] = {nullptr, nullptr};
  int pidx = 0;             // p64 buffer holding the last projection's pressure
  bool have_p = false;      // a previous pressure exists (warm start)
  std::vector<Level> lv;    // multigrid hierarchy; lv[0].x = e, lv[0].rhs = r
  double* scal = nullptr;   // red[4] followed by rmax[rmax_len], cleared by one fill
  double* red = nullptr;    // scal + 0: [0] max|u|, [1] max|v|, [2] dt
  double* rmax = nullptr;   // scal + 4: per-iteration FP64 residual maxima
  int rmax_len = 0;
  int p32idx = 0;           // p32 buffer holding the last flat-solver pressure
  double* diag_part = nullptr;
  double* diag_out = nullptr;
  int diag_blocks = 0;
  double* host = nullptr;   // pinned staging for scalars
  bool have_maxabs = false; // red[0..1] hold max|u|, max|v| of the current field
  // forcing
  bool forcing_on = false;
  fs_forcing forcing{};
  // stats
  double t = 0;
  int64_t steps = 0;
  int last_vcycles = 0;
  int64_t total_vcycles = 0;
  double last_residual = 0;
  double last_dt = 0;
  int last_pred = 2;        // speculative V-cycle count for the next projection
  Profile prof;
};

namespace {

inline dim3 grid2(int nx, int ny) {
  return dim3((unsigned)((nx + fsk::BX - 1) / fsk::BX), (unsigned)((ny + fsk::BY - 1) / fsk::BY));
}
inline dim3 grid_pairs(int nx, int ny) {
  int half = (nx + 1) / 2;
  return dim3((unsigned)((half + fsk::BX - 1) / fsk::BX), (unsigned)((ny + fsk::BY - 1) / fsk::BY));
}
const dim3 BLK(fsk::BX, fsk::BY);

template <typename T>
int dalloc(T** p, size_t count) {
  hipError_t e = hipMalloc((void**)p, count * sizeof(T) > 0 ? count * sizeof(T) : 1);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? FS_E_OOM : FS_E_DEVICE, "hipMalloc(%zu bytes) failed: %s",
                count * sizeof(T), hipGetErrorString(e));
  }
  return FS_OK;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}

// event pair around launches of the profiled kernel class
struct ProfScope {
  fs_solver_t* s;
  bool on;
  hipEvent_t a{}, b{};
  ProfScope(fs_solver_t* s_, int kind) : s(s_), on(s_->prof.kind == kind) {
    if (!on) return;
    a = take();
    b = take();
    (void)hipEventRecord(a, s->stream);
  }
  hipEvent_t take() {
    if (!s->prof.pool.empty()) {
      hipEvent_t e = s->prof.pool.back();
      s->prof.pool.pop_back();
      return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(b, s->stream);
    s->prof.pending.push_back({a, b});
  }
};

void prof_collect(fs_solver_t* s) {
  for (auto& pr : s->prof.pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
      s->prof.total_ms += ms;
      s->prof.launches++;
    }
    s->prof.pool.push_back(pr.first);
    s->prof.pool.push_back(pr.second);
  }
  s->prof.pending.clear();
}

int check_solver(void* h) {
  if (!h) return fail(FS_E_INVALID_ARG, "solver handle is NULL");
  return FS_OK;
}

int set_device(fs_solver_t* s) {
  HIPCHK(hipSetDevice(s->dev));
  return FS_OK;
}

// ------------------------------------------------------------- step pieces

int enqueue_advection(fs_solver_t* s, double dt) {
  ProfScope ps(s, 0);
  float cx = (float)(dt / s->dx), cy = (float)(dt / s->dy);
  int c = s->cur;
  if (s->cfg.advection == FS_ADV_MACCORMACK) {
    fsk::k_advect_sl<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->fu, s->fv, s->nx, s->ny,
                                                                   cx, cy);
    LAUNCHCHK();
    fsk::k_advect_mc<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->fu, s->fv, s->u[1 - c],
                                                                   s->v[1 - c], s->nx, s->ny, cx, cy);
    LAUNCHCHK();
  } else {
    fsk::k_advect_sl<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->u[1 - c], s->v[1 - c],
                                                                   s->nx, s->ny, cx, cy);
    LAUNCHCHK();
  }
  s->cur = 1 - c;
  return FS_OK;
}

int enqueue_diffusion(fs_solver_t* s, double dt) {
  if (!(s->cfg.nu > 0)) return FS_OK;
  float a = (float)(s->cfg.nu * dt / (s->dx * s->dx)), b = (float)(s->cfg.nu * dt / (s->dy * s->dy));
  int c = s->cur;
  fsk::k_diffuse<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->u[1 - c], s->v[1 - c], s->nx,
                                                               s->ny, a, b);
  LAUNCHCHK();
  s->cur = 1 - c;
  return FS_OK;
}

double js_round(double x) {
  double r = std::floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}

// 24 modes of src/cpu/solver.js:181-199 from the exact uint32 LCG.
int enqueue_forcing(fs_solver_t* s, double dt) {
  if (!s->forcing_on) return FS_OK;
  const fs_forcing& f = s->forcing;
  uint32_t st = f.seed + (uint32_t)s->steps;
  auto rnd = [&]() {
    st = st * 1664525u + 1013904223u;
    return (double)st / 4294967296.0;
  };
  fsk::ModeSet ms{};
  const int M = 24;
  double scale = f.amp * dt / std::sqrt((double)M);
  int k = 0;
  for (int m = 0; m < M; m++) {
    double theta = 2 * M_PI * rnd();
    double kmag = f.kf + (rnd() - 0.5) * 2 * f.width;
    int kx = (int)js_round(kmag * std::cos(theta));
    int ky = (int)js_round(kmag * std::sin(theta));
    double ph = 2 * M_PI * rnd();
    if (kx == 0 && ky == 0) continue;
    double kk = std::sqrt((double)kx * kx + (double)ky * ky);
    fsk::Mode& md = ms.m[k++];
    md.tx = kx * s->dx / (2 * M_PI);
    md.ty = ky * s->dy / (2 * M_PI);
    md.tph = ph / (2 * M_PI);
    md.cu = (float)(scale * (ky / kk));
    md.cv = (float)(scale * (-kx / kk));
  }
  ms.n = k;
  float drag = f.drag > 0 ? (float)std::exp(-f.drag * dt) : 1.0f;
  int c = s->cur;
  fsk::k_forcing<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->nx, s->ny, ms, drag);
  LAUNCHCHK();
  return FS_OK;
}

int fill_f32(fs_solver_t* s, float* a, size_t n, float val) {
  fsk::k_fill_f32<<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(a, n, val);
  LAUNCHCHK();
  return FS_OK;
}
int fill_f64(fs_solver_t* s, double* a, size_t n, double val) {
  fsk::k_fill_f64<<<(unsigned)((n + 255) / 256), 256, 0, s->stream>>>(a, n, val);
  LAUNCHCHK();
  return FS_OK;
}

int enqueue_divergence(fs_solver_t* s) {
  int c = s->cur;
  fsk::k_divergence<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->div, s->nx, s->ny,
                                                                  (float)(1.0 / s->dx), (float)(1.0 / s->dy));
  LAUNCHCHK();
  return FS_OK;
}

// Jacobi / red-black with FP32 pressure; result in p32[result index], returns index via out.
int solve_flat(fs_solver_t* s, int iterations, int* out_idx) {
  float h2 = (float)(s->dx * s->dx);
  int rc = fill_f32(s, s->p32[0], s->n, 0.f);
  if (rc) return rc;
  int a = 0;
  if (s->cfg.pressure_solver == FS_P_JACOBI) {
    for (int k = 0; k < iterations; k++) {
      fsk::k_jacobi<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->p32[a], s->div, s->p32[1 - a], s->nx, s->ny,
                                                                  h2, 1.0f);
      LAUNCHCHK();
      a = 1 - a;
    }
  } else {
    for (int k = 0; k < iterations; k++) {
      for (int colour = 0; colour < 2; colour++) {
        ProfScope ps(s, 1);
        fsk::k_rb_colour<float><<<grid_pairs(s->nx, s->ny), BLK, 0, s->stream>>>(s->p32[0], s->div, s->nx, s->ny,
                                                                                  h2, 1.0f, colour, 0);
        LAUNCHCHK();
      }
    }
  }
  *out_idx = a;
  return FS_OK;
}

// Smooth `sweeps` red-black sweeps on level l; x_zero marks a logically-zero start.
int mg_smooth(fs_solver_t* s, int l, int sweeps, bool& x_zero, const double* conv, double tol) {
  Level& L = s->lv[l];
  float h2 = (float)(L.dx * L.dx);
  for (int k = 0; k < sweeps; k++) {
    for (int colour = 0; colour < 2; colour++) {
      ProfScope ps(s, l == 0 ? 1 : -2);
      fsk::k_mg_rb<float><<<grid_pairs(L.n, L.n), BLK, 0, s->stream>>>(L.x, L.rhs, L.n, h2, colour, x_zero ? 1 : 0,
                                                                        conv, tol);
      LAUNCHCHK();
      // after the red pass from zero, black cells are still stale but the black
      // pass never reads its own value, and red cells are now valid
      x_zero = false;
    }
  }
  return FS_OK;
}

// One V-cycle on the correction (lv[0].x = e, lv[0].rhs = r), predicated on conv.
int mg_vcycle(fs_solver_t* s, const double* conv, double tol) {
  int Lc = (int)s->lv.size();
  std::vector<char> zero(Lc, 1);  // every level's correction starts from zero
  for (int l = 0; l < Lc - 1; l++) {
    bool z = zero[l];
    int rc = mg_smooth(s, l, s->cfg.mg_pre, z, conv, tol);
    if (rc) return rc;
    zero[l] = z;
    Level& F = s->lv[l];
    Level& Cl = s->lv[l + 1];
    fsk::k_mg_restrict<float><<<grid2(Cl.n, Cl.n), BLK, 0, s->stream>>>(
        F.x, F.rhs, F.n, (float)(1.0 / (F.dx * F.dx)), Cl.rhs, zero[l] ? 1 : 0, conv, tol);
    LAUNCHCHK();
  }
  {
    Level& C = s->lv[Lc - 1];
    float h2 = (float)(C.dx * C.dx);
    if (C.n <= fsk::COARSE_MAX) {
      fsk::k_mg_coarse<float><<<1, 256, 0, s->stream>>>(C.x, C.rhs, C.n, h2, s->cfg.mg_coarse, conv, tol);
      LAUNCHCHK();
    } else {
      bool z = true;
      int rc = mg_smooth(s, Lc - 1, s->cfg.mg_coarse, z, conv, tol);
      if (rc) return rc;
      if (z) {  // zero sweeps: the coarse correction is zero
        int rc2 = fill_f32(s, C.x, (size_t)C.n * C.n, 0.f);
        if (rc2) return rc2;
      }
    }
  }
  for (int l = Lc - 2; l >= 0; l--) {
    Level& F = s->lv[l];
    Level& Cl = s->lv[l + 1];
    fsk::k_mg_prolong<float><<<grid2(F.n, F.n), BLK, 0, s->stream>>>(Cl.x, F.x, F.n, zero[l] ? 1 : 0, conv, tol);
    LAUNCHCHK();
    bool z = false;
    int rc = mg_smooth(s, l, s->cfg.mg_post, z, conv, tol);
    if (rc) return rc;
  }
  return FS_OK;
}

// FP64 residual of iteration `it`: reads p64[pin] (+ e if use_e), writes p64[1-pin], r.
int enqueue_residual(fs_solver_t* s, int it, int pin, bool zero_p, bool use_e, double tol) {
  ProfScope ps(s, 2);
  int c = s->cur;
  const double* skip = it > 0 ? s->rmax + (it - 1) : nullptr;
  double idx = 1.0 / s->dx, idy = 1.0 / s->dy;
  fsk::k_residual64<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(
      s->u[c], s->v[c], zero_p ? nullptr : s->p64[pin], use_e ? s->lv[0].x : nullptr, s->p64[1 - pin],
      s->lv[0].rhs, s->nx, s->ny, idx, idy, idx * idx, idy * idy, s->rmax + it, skip, tol);
  LAUNCHCHK();
  return FS_OK;
}

// Multigrid projection by iterative refinement. Returns the V-cycle count via *cycles
// and leaves the final pressure in p64[s->pidx].
int solve_multigrid(fs_solver_t* s, int* cycles_out, double* resid_out) {
  const double tol_cfg = s->cfg.tolerance;
  const bool tol_mode = tol_cfg > 0;
  const double tol = tol_mode ? tol_cfg : -1.0;  // -1: never converged (fixed count)
  const int maxc = std::max(0, s->cfg.pressure_iterations);
  // clears max|u|, max|v| for the gradient kernel and every residual slot
  int rc = fill_f64(s, s->scal, (size_t)s->rmax_len + 4, 0.0);
  if (rc) return rc;
  bool zero_p = !(s->cfg.guess == FS_GUESS_PREVIOUS && s->have_p);
  int a0 = s->pidx;
  int it = 0;
  int cycles = 0;
  if (!tol_mode) {
    // k V-cycles: residual(0), V(0), residual(1), ..., V(k-1), residual(k)
    for (it = 0; it <= maxc; it++) {
      int pin = (a0 + it) & 1;
      rc = enqueue_residual(s, it, pin, it == 0 && zero_p, it > 0, tol);
      if (rc) return rc;
      if (it < maxc) {
        rc = mg_vcycle(s, s->rmax + it, tol);
        if (rc) return rc;
      }
    }
    cycles = maxc;
    s->pidx = (a0 + maxc + 1) & 1;
    HIPCHK(hipMemcpyAsync(s->host, s->rmax + maxc, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    prof_collect(s);
    *resid_out = s->host[0];
  } else {
    // speculate last_pred iterations, then check; continue one at a time
    int done_at = -1;
    int batch = std::max(1, std::min(s->last_pred, maxc + 1));
    while (done_at < 0) {
      int start = it;
      for (; it < start + batch && it <= maxc; it++) {
        int pin = (a0 + it) & 1;
        rc = enqueue_residual(s, it, pin, it == 0 && zero_p, it > 0, tol);
        if (rc) return rc;
        if (it < maxc) {
          rc = mg_vcycle(s, s->rmax + it, tol);
          if (rc) return rc;
        }
      }
      HIPCHK(hipMemcpyAsync(s->host, s->rmax, sizeof(double) * it, hipMemcpyDeviceToHost, s->stream));
      HIPCHK(hipStreamSynchronize(s->stream));
      prof_collect(s);
      for (int k = 0; k < it; k++) {
        double r = s->host[k];
        if (std::isnan(r) || std::isinf(r)) {
          return fail(FS_E_NONFINITE, "non-finite residual in projection (iteration %d)", k);
        }
        if (r <= tol) {
          done_at = k;
          break;
        }
      }
      if (done_at < 0 && it > maxc) {
        s->pidx = (a0 + maxc + 1) & 1;
        s->last_vcycles = maxc;
        *cycles_out = maxc;
        *resid_out = s->host[maxc];
        s->last_pred = maxc + 1;
        return fail(FS_E_NOT_CONVERGED, "multigrid did not reach tolerance %.3g in %d V-cycles (residual %.3g)",
                    tol_cfg, maxc, s->host[maxc]);
      }
      batch = 1;
    }
    cycles = done_at;
    s->pidx = (a0 + done_at + 1) & 1;
    *resid_out = s->host[done_at];
    s->last_pred = done_at + 1;
  }
  *cycles_out = cycles;
  return FS_OK;
}

int enqueue_projection(fs_solver_t* s, int* cycles, double* resid) {
  *cycles = 0;
  *resid = 0;
  int c = s->cur;
  if (s->cfg.pressure_solver == FS_P_MULTIGRID) {
    int rc = solve_multigrid(s, cycles, resid);
    if (rc && rc != FS_E_NOT_CONVERGED) return rc;
    s->have_p = true;
    {
      ProfScope ps(s, 3);
      double idx = 1.0 / s->dx, idy = 1.0 / s->dy;
      fsk::k_gradient64<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->p64[s->pidx],
                                                                      s->u[1 - c], s->v[1 - c], s->nx, s->ny, idx,
                                                                      idy, s->red);
      LAUNCHCHK();
    }
    s->have_maxabs = true;
    s->cur = 1 - c;
    return rc;
  }
  int rc = enqueue_divergence(s);
  if (rc) return rc;
  int pi = 0;
  rc = solve_flat(s, s->cfg.pressure_iterations, &pi);
  if (rc) return rc;
  s->p32idx = pi;
  {
    ProfScope ps(s, 3);
    fsk::k_gradient32<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->p32[pi], s->u[1 - c],
                                                                    s->v[1 - c], s->nx, s->ny,
                                                                    (float)(1.0 / s->dx), (float)(1.0 / s->dy));
    LAUNCHCHK();
  }
  s->have_maxabs = false;
  s->cur = 1 - c;
  return FS_OK;
}

int do_step(fs_solver_t* s, double dt) {
  int rc = set_device(s);
  if (rc) return rc;
  if (!(dt >= 0) || !std::isfinite(dt)) return fail(FS_E_INVALID_ARG, "dt must be finite and >= 0, got %g", dt);
  rc = enqueue_advection(s, dt);
  if (rc) return rc;
  rc = enqueue_diffusion(s, dt);
  if (rc) return rc;
  rc = enqueue_forcing(s, dt);
  if (rc) return rc;
  int cycles = 0;
  double resid = 0;
  int prc = enqueue_projection(s, &cycles, &resid);
  if (prc && prc != FS_E_NOT_CONVERGED) return prc;
  s->last_vcycles = cycles;
  s->total_vcycles += cycles;
  s->last_residual = resid;
  s->last_dt = dt;
  s->t += dt;
  s->steps++;
  return prc;
}

int build_levels(fs_solver_t* s) {
  if (s->cfg.pressure_solver != FS_P_MULTIGRID) return FS_OK;
  if (s->nx != s->ny) return fail(FS_E_INVALID_ARG, "multigrid here assumes a square grid, got %dx%d", s->nx, s->ny);
  int minS = s->cfg.mg_min_size;
  if (minS < 2) return fail(FS_E_INVALID_ARG, "mg_min_size must be >= 2, got %d", minS);
  int n = s->nx;
  if (n < minS) return fail(FS_E_INVALID_ARG, "grid %d is smaller than mg_min_size %d", n, minS);
  double L = s->cfg.lx;
  // src/gpu/multigrid.js:182-197: halve while n >= minSize (and n stays even)
  while (n >= minS) {
    Level lvl;
    lvl.n = n;
    lvl.dx = L / n;
    s->lv.push_back(lvl);
    if (n / 2 < minS || (n & 1)) break;
    n /= 2;
  }
  for (size_t l = 0; l < s->lv.size(); l++) {
    size_t c = (size_t)s->lv[l].n * s->lv[l].n;
    int rc = dalloc(&s->lv[l].x, c);
    if (rc) return rc;
    rc = dalloc(&s->lv[l].rhs, c);
    if (rc) return rc;
  }
  return FS_OK;
}

void release(fs_solver_t* s) {
  if (!s) return;
  (void)hipSetDevice(s->dev);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (int k = 0; k < 2; k++) {
    dfree(s->u[k]);
    dfree(s->v[k]);
    dfree(s->p32[k]);
    dfree(s->p64[k]);
  }
  dfree(s->fu);
  dfree(s->fv);
  dfree(s->div);
  for (auto& l : s->lv) {
    dfree(l.x);
    dfree(l.rhs);
  }
  dfree(s->scal);
  dfree(s->diag_part);
  dfree(s->diag_out);
  if (s->host) (void)hipHostFree(s->host);
  for (auto& pr : s->prof.pending) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  for (auto e : s->prof.pool) (void)hipEventDestroy(e);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int run_diag(fs_solver_t* s, double out[8]) {
  int c = s->cur;
  dim3 g = grid2(s->nx, s->ny);
  fsk::k_diag<<<g, BLK, 0, s->stream>>>(s->u[c], s->v[c], s->nx, s->ny, 1.0 / s->dx, 1.0 / s->dy, s->diag_part);
  LAUNCHCHK();
  fsk::k_diag_final<<<1, 256, 0, s->stream>>>(s->diag_part, (int)(g.x * g.y), s->diag_out);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(s->host, s->diag_out, sizeof(double) * 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  prof_collect(s);
  for (int k = 0; k < 8; k++) out[k] = s->host[k];
  return FS_OK;
}

// grow the residual-slot buffer (and its pinned mirror) to hold `iterations` + 2 slots
int ensure_rmax(fs_solver_t* s, int iterations) {
  if (iterations + 2 <= s->rmax_len) return FS_OK;
  int rc = set_device(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  double* nr = nullptr;
  if ((rc = dalloc(&nr, (size_t)iterations + 2 + 4))) return rc;
  if (hipMemcpy(nr, s->scal, 4 * sizeof(double), hipMemcpyDeviceToDevice) != hipSuccess) {
    dfree(nr);
    return fail(FS_E_DEVICE, "copy of the scalar block failed");
  }
  double* nh = nullptr;
  if (hipHostMalloc((void**)&nh, ((size_t)iterations + 10) * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    dfree(nr);
    return fail(FS_E_OOM, "hipHostMalloc failed");
  }
  dfree(s->scal);
  (void)hipHostFree(s->host);
  s->scal = nr;
  s->red = nr;
  s->rmax = nr + 4;
  s->host = nh;
  s->rmax_len = iterations + 2;
  return FS_OK;
}

}  // namespace

// ======================================================================= ABI

extern "C" {

int fs_abi_version(void) { return FS_ABI_VERSION; }

const char* fs_last_error(void) { return g_err.c_str(); }

int fs_device_count(int* count) {
  if (!count) return fail(FS_E_INVALID_ARG, "count is NULL");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  if (e != hipSuccess) {
    *count = 0;
    return fail(FS_E_DEVICE, "hipGetDeviceCount failed: %s", hipGetErrorString(e));
  }
  *count = c;
  return FS_OK;
}

int fs_device_name(int device, char* buf, size_t len) {
  if (!buf || len == 0) return fail(FS_E_INVALID_ARG, "buffer is NULL");
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device));
  snprintf(buf, len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return FS_OK;
}

void fs_config_default(fs_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof *c);
  c->nx = c->ny = 256;
  c->lx = c->ly = 2 * M_PI;
  c->nu = 0.0;
  c->advection = FS_ADV_MACCORMACK;
  c->pressure_solver = FS_P_JACOBI;
  c->pressure_iterations = 40;
  c->tolerance = 0.0;
  c->mg_min_size = 8;
  c->mg_pre = 2;
  c->mg_post = 2;
  c->mg_coarse = 40;
  c->mg_precision = FS_MG_FP32;
  c->guess = FS_GUESS_ZERO;
  c->solid_mask_width = 1.5;
  c->device = 0;
}

int fs_solver_create(const fs_config* cfg, void** out) {
  if (!cfg || !out) return fail(FS_E_INVALID_ARG, "config or output pointer is NULL");
  *out = nullptr;
  const fs_config& c = *cfg;
  if (c.nx < 2 || c.ny < 2) return fail(FS_E_INVALID_ARG, "grid must be >= 2x2, got %dx%d", c.nx, c.ny);
  if ((int64_t)c.nx * c.ny > (int64_t)1 << 28)
    return fail(FS_E_INVALID_ARG, "grid %dx%d exceeds 2^28 cells", c.nx, c.ny);
  if (!(c.lx > 0) || !(c.ly > 0)) return fail(FS_E_INVALID_ARG, "domain size must be positive");
  if (c.advection != FS_ADV_SEMI_LAGRANGIAN && c.advection != FS_ADV_MACCORMACK)
    return fail(FS_E_INVALID_ARG, "unknown advection scheme %d; have semi-lagrangian(0), maccormack(1)", c.advection);
  if (c.pressure_solver != FS_P_JACOBI && c.pressure_solver != FS_P_RED_BLACK && c.pressure_solver != FS_P_MULTIGRID)
    return fail(FS_E_INVALID_ARG, "unknown pressure solver %d; have jacobi(0), red-black(1), multigrid(2)",
                c.pressure_solver);
  if (c.pressure_iterations < 0) return fail(FS_E_INVALID_ARG, "pressure_iterations must be >= 0");
  if (c.tolerance < 0 || !std::isfinite(c.tolerance)) return fail(FS_E_INVALID_ARG, "tolerance must be >= 0");
  if (c.tolerance > 0 && c.pressure_solver != FS_P_MULTIGRID)
    return fail(FS_E_INVALID_ARG, "a residual tolerance needs the multigrid solver");
  if (c.mg_pre < 0 || c.mg_post < 0 || c.mg_coarse < 0) return fail(FS_E_INVALID_ARG, "negative sweep count");
  if (c.mg_precision != FS_MG_FP32)
    return fail(FS_E_UNSUPPORTED, "multigrid storage precision %d is not built yet", c.mg_precision);
  if (c.guess != FS_GUESS_ZERO && c.guess != FS_GUESS_PREVIOUS)
    return fail(c.guess == FS_GUESS_LEARNED ? FS_E_UNSUPPORTED : FS_E_INVALID_ARG, "guess mode %d not available",
                c.guess);
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(FS_E_DEVICE, "no HIP device available");
  if (c.device < 0 || c.device >= ndev) return fail(FS_E_INVALID_ARG, "device %d out of range (%d)", c.device, ndev);

  fs_solver_t* s = new fs_solver_t();
  s->cfg = c;
  s->dev = c.device;
  s->nx = c.nx;
  s->ny = c.ny;
  s->n = (size_t)c.nx * c.ny;
  s->dx = c.lx / c.nx;
  s->dy = c.ly / c.ny;
  int rc = FS_OK;
  auto bail = [&](int code) {
    std::string msg = g_err;
    release(s);
    g_err = msg;
    return code;
  };
  if (hipSetDevice(s->dev) != hipSuccess) return bail(fail(FS_E_DEVICE, "hipSetDevice(%d) failed", s->dev));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FS_E_DEVICE, "hipStreamCreate failed"));
  for (int k = 0; k < 2 && !rc; k++) {
    rc = dalloc(&s->u[k], s->n);
    if (!rc) rc = dalloc(&s->v[k], s->n);
  }
  if (!rc && c.advection == FS_ADV_MACCORMACK) {
    rc = dalloc(&s->fu, s->n);
    if (!rc) rc = dalloc(&s->fv, s->n);
  }
  if (!rc) rc = dalloc(&s->div, s->n);
  if (!rc && c.pressure_solver != FS_P_MULTIGRID) {
    rc = dalloc(&s->p32[0], s->n);
    if (!rc) rc = dalloc(&s->p32[1], s->n);
  }
  if (!rc) rc = dalloc(&s->p64[0], s->n);
  if (!rc) rc = dalloc(&s->p64[1], s->n);
  if (!rc) rc = build_levels(s);
  s->rmax_len = std::max(c.pressure_iterations, 1) + 2;
  if (!rc) rc = dalloc(&s->scal, (size_t)s->rmax_len + 4);
  if (!rc) {
    s->red = s->scal;
    s->rmax = s->scal + 4;
  }
  dim3 g = grid2(s->nx, s->ny);
  s->diag_blocks = (int)(g.x * g.y);
  if (!rc) rc = dalloc(&s->diag_part, (size_t)s->diag_blocks * fsk::DIAG_W);
  if (!rc) rc = dalloc(&s->diag_out, 8);
  if (!rc) {
    size_t hn = std::max<size_t>(64, (size_t)s->rmax_len + 8);
    if (hipHostMalloc((void**)&s->host, hn * sizeof(double), hipHostMallocDefault) != hipSuccess)
      rc = fail(FS_E_OOM, "hipHostMalloc of the staging buffer failed");
  }
  if (rc) return bail(rc);
  // zero state: velocity, pressure, accumulators
  if (!rc) rc = fill_f32(s, s->u[0], s->n, 0.f);
  if (!rc) rc = fill_f32(s, s->v[0], s->n, 0.f);
  if (!rc) rc = fill_f64(s, s->p64[0], s->n, 0.0);
  if (!rc) rc = fill_f64(s, s->p64[1], s->n, 0.0);
  if (!rc) rc = fill_f32(s, s->div, s->n, 0.f);
  if (!rc) rc = fill_f64(s, s->scal, (size_t)s->rmax_len + 4, 0.0);
  for (auto& l : s->lv) {
    if (!rc) rc = fill_f32(s, l.x, (size_t)l.n * l.n, 0.f);
    if (!rc) rc = fill_f32(s, l.rhs, (size_t)l.n * l.n, 0.f);
  }
  if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = fail(FS_E_DEVICE, "initial fill failed");
  if (rc) return bail(rc);
  *out = s;
  return FS_OK;
}

int fs_solver_destroy(void* h) {
  release((fs_solver_t*)h);
  return FS_OK;
}

int fs_upload_velocity_f32(void* h, const float* u, const float* v, size_t n) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!u || !v) return fail(FS_E_INVALID_ARG, "u or v is NULL");
  if (n != s->n) return fail(FS_E_INVALID_ARG, "expected %zu values per component, got %zu", s->n, n);
  if ((rc = set_device(s))) return rc;
  HIPCHK(hipMemcpyAsync(s->u[s->cur], u, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->v[s->cur], v, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->have_maxabs = false;
  s->have_p = false;
  return FS_OK;
}

int fs_upload_velocity_f64(void* h, const double* u, const double* v, size_t n) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!u || !v) return fail(FS_E_INVALID_ARG, "u or v is NULL");
  if (n != s->n) return fail(FS_E_INVALID_ARG, "expected %zu values per component, got %zu", s->n, n);
  std::vector<float> fu(n), fv(n);
  for (size_t k = 0; k < n; k++) {
    fu[k] = (float)u[k];
    fv[k] = (float)v[k];
  }
  return fs_upload_velocity_f32(h, fu.data(), fv.data(), n);
}

int fs_read_velocity(void* h, float* u, float* v, size_t n) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!u || !v) return fail(FS_E_INVALID_ARG, "u or v is NULL");
  if (n != s->n) return fail(FS_E_INVALID_ARG, "expected %zu values per component, got %zu", s->n, n);
  if ((rc = set_device(s))) return rc;
  HIPCHK(hipMemcpyAsync(u, s->u[s->cur], n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipMemcpyAsync(v, s->v[s->cur], n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  prof_collect(s);
  return FS_OK;
}

int fs_read_field(void* h, int field, void* out, size_t bytes) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!out) return fail(FS_E_INVALID_ARG, "output is NULL");
  const void* src = nullptr;
  size_t need = 0;
  switch (field) {
    case FS_F_U: src = s->u[s->cur]; need = s->n * sizeof(float); break;
    case FS_F_V: src = s->v[s->cur]; need = s->n * sizeof(float); break;
    case FS_F_DIV: src = s->div; need = s->n * sizeof(float); break;
    case FS_F_PRESSURE:
      need = s->n * sizeof(double);
      if (s->cfg.pressure_solver != FS_P_MULTIGRID) {
        if (bytes != need) return fail(FS_E_INVALID_ARG, "field %d needs %zu bytes, got %zu", field, need, bytes);
        if ((rc = set_device(s))) return rc;
        std::vector<float> pf(s->n);
        HIPCHK(hipMemcpyAsync(pf.data(), s->p32[s->p32idx], s->n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        for (size_t k = 0; k < s->n; k++) ((double*)out)[k] = pf[k];
        return FS_OK;
      }
      src = s->p64[s->pidx];
      break;
    case FS_F_RESIDUAL:
      if (s->lv.empty()) return fail(FS_E_INVALID_ARG, "no multigrid residual without the multigrid solver");
      src = s->lv[0].rhs; need = s->n * sizeof(float); break;
    default: return fail(FS_E_INVALID_ARG, "unknown field %d", field);
  }
  if (bytes != need) return fail(FS_E_INVALID_ARG, "field %d needs %zu bytes, got %zu", field, need, bytes);
  if ((rc = set_device(s))) return rc;
  HIPCHK(hipMemcpyAsync(out, src, need, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return FS_OK;
}

int fs_step(void* h, double dt) {
  int rc = check_solver(h);
  if (rc) return rc;
  return do_step((fs_solver_t*)h, dt);
}

int fs_step_cfl(void* h, double cfl) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!(cfl > 0) || !std::isfinite(cfl)) return fail(FS_E_INVALID_ARG, "cfl must be > 0");
  if ((rc = set_device(s))) return rc;
  if (!s->have_maxabs) {
    fill_f64(s, s->red, 2, 0.0);
    fsk::k_maxabs<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[s->cur], s->v[s->cur], s->nx, s->ny, s->red);
    LAUNCHCHK();
  }
  // dt is needed on the host for the kernel arguments of this step
  fsk::k_cfl_dt<<<1, 1, 0, s->stream>>>(s->red, cfl, 1.0 / s->dx, 1.0 / s->dy, 1e30, s->red + 2);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(s->host, s->red + 2, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  double dt = s->host[0];
  if (!std::isfinite(dt) || dt >= 1e30) return fail(FS_E_NONFINITE, "CFL time step undefined (zero or non-finite velocity)");
  return do_step(s, dt);
}

int fs_sync(void* h) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  prof_collect(s);
  return FS_OK;
}

int fs_max_divergence(void* h, double* out) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!out) return fail(FS_E_INVALID_ARG, "out is NULL");
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  if ((rc = enqueue_divergence(s))) return rc;  // the reference also refreshes its div texture
  double d[8];
  if ((rc = run_diag(s, d))) return rc;
  *out = d[1];
  return FS_OK;
}

int fs_poisson_residual_after(void* h, int iterations, double* out) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!out) return fail(FS_E_INVALID_ARG, "out is NULL");
  if (iterations < 0) return fail(FS_E_INVALID_ARG, "iterations must be >= 0");
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  // src/gpu/solver.js:857-896: divergence, zero p, `iterations` of the solver, then
  // ||L p - div|| / ||div|| evaluated on the host in f64 from the two fields.
  if ((rc = enqueue_divergence(s))) return rc;
  std::vector<float> dv(s->n);
  std::vector<double> p(s->n);
  if (s->cfg.pressure_solver == FS_P_MULTIGRID) {
    if ((rc = ensure_rmax(s, iterations))) return rc;
    fs_config saved = s->cfg;
    s->cfg.tolerance = 0;
    s->cfg.pressure_iterations = iterations;
    s->cfg.guess = FS_GUESS_ZERO;
    int cyc = 0;
    double res = 0;
    rc = solve_multigrid(s, &cyc, &res);
    s->cfg = saved;
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(p.data(), s->p64[s->pidx], s->n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  } else {
    int pi = 0;
    if ((rc = solve_flat(s, iterations, &pi))) return rc;
    std::vector<float> pf(s->n);
    HIPCHK(hipMemcpyAsync(pf.data(), s->p32[pi], s->n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (size_t k = 0; k < s->n; k++) p[k] = pf[k];
  }
  HIPCHK(hipMemcpyAsync(dv.data(), s->div, s->n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  prof_collect(s);
  int nx = s->nx, ny = s->ny;
  double h2 = s->dx * s->dx;
  double num = 0, den = 0;
  for (int j = 0; j < ny; j++) {
    int jm = j == 0 ? ny - 1 : j - 1, jp = j + 1 == ny ? 0 : j + 1;
    for (int i = 0; i < nx; i++) {
      int im = i == 0 ? nx - 1 : i - 1, ip = i + 1 == nx ? 0 : i + 1;
      double lap = (p[j * nx + im] + p[j * nx + ip] + p[jm * nx + i] + p[jp * nx + i] - 4 * p[j * nx + i]) / h2;
      double r = lap - dv[j * nx + i];
      num += r * r;
      den += (double)dv[j * nx + i] * dv[j * nx + i];
    }
  }
  *out = std::sqrt(num) / (den > 0 ? std::sqrt(den) : 1.0);
  return FS_OK;
}

int fs_set_pressure_iterations(void* h, int iterations) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (iterations < 0) return fail(FS_E_INVALID_ARG, "iterations must be >= 0");
  if ((rc = ensure_rmax(s, iterations))) return rc;
  s->cfg.pressure_iterations = iterations;
  return FS_OK;
}

int fs_set_tolerance(void* h, double tol) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (tol < 0 || !std::isfinite(tol)) return fail(FS_E_INVALID_ARG, "tolerance must be >= 0");
  if (tol > 0 && s->cfg.pressure_solver != FS_P_MULTIGRID)
    return fail(FS_E_INVALID_ARG, "a residual tolerance needs the multigrid solver");
  s->cfg.tolerance = tol;
  return FS_OK;
}

int fs_set_nu(void* h, double nu) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!(nu >= 0) || !std::isfinite(nu)) return fail(FS_E_INVALID_ARG, "nu must be >= 0");
  ((fs_solver_t*)h)->cfg.nu = nu;
  return FS_OK;
}

int fs_set_guess(void* h, int guess) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (guess != FS_GUESS_ZERO && guess != FS_GUESS_PREVIOUS)
    return fail(guess == FS_GUESS_LEARNED ? FS_E_UNSUPPORTED : FS_E_INVALID_ARG, "guess mode %d not available", guess);
  ((fs_solver_t*)h)->cfg.guess = guess;
  return FS_OK;
}

int fs_set_forcing(void* h, const fs_forcing* f) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!f) {
    s->forcing_on = false;
    return FS_OK;
  }
  if (!std::isfinite(f->kf) || !std::isfinite(f->width) || !std::isfinite(f->amp) || !(f->drag >= 0))
    return fail(FS_E_INVALID_ARG, "forcing parameters must be finite, drag >= 0");
  s->forcing = *f;
  s->forcing_on = true;
  return FS_OK;
}

int fs_get_stats(void* h, fs_stats* out) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!out) return fail(FS_E_INVALID_ARG, "out is NULL");
  fs_solver_t* s = (fs_solver_t*)h;
  out->t = s->t;
  out->steps = s->steps;
  out->last_vcycles = s->last_vcycles;
  out->total_vcycles = s->total_vcycles;
  out->last_residual = s->last_residual;
  out->last_dt = s->last_dt;
  return FS_OK;
}

int fs_diagnostics(void* h, fs_diag* out) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!out) return fail(FS_E_INVALID_ARG, "out is NULL");
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  double d[8];
  if ((rc = run_diag(s, d))) return rc;
  double N = (double)s->n, dA = s->dx * s->dy;
  out->kinetic_energy = 0.5 * d[0] / N;
  out->max_divergence = d[1];
  out->l2_divergence = std::sqrt(d[2] / N);
  out->momentum_x = d[3] * dA;
  out->momentum_y = d[4] * dA;
  out->max_abs_u = d[5];
  out->max_abs_v = d[6];
  out->nonfinite = (int)d[7];
  if (out->nonfinite) return fail(FS_E_NONFINITE, "%d non-finite velocity faces", out->nonfinite);
  return FS_OK;
}

int fs_run_kernel(void* h, int kernel, double dt) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  switch (kernel) {
    case FS_K_ADVECT_SL: {
      int saved = s->cfg.advection;
      s->cfg.advection = FS_ADV_SEMI_LAGRANGIAN;
      rc = enqueue_advection(s, dt);
      s->cfg.advection = saved;
      break;
    }
    case FS_K_ADVECT_MC:
      if (s->cfg.advection != FS_ADV_MACCORMACK)
        return fail(FS_E_INVALID_ARG, "MacCormack scratch exists only on a maccormack solver");
      rc = enqueue_advection(s, dt);
      break;
    case FS_K_DIVERGENCE: rc = enqueue_divergence(s); break;
    case FS_K_DIFFUSE: rc = enqueue_diffusion(s, dt); break;
    case FS_K_PROJECT: {
      int cyc = 0;
      double res = 0;
      rc = enqueue_projection(s, &cyc, &res);
      s->last_vcycles = cyc;
      s->last_residual = res;
      break;
    }
    case FS_K_GRADIENT: {
      int c = s->cur;
      double idx = 1.0 / s->dx, idy = 1.0 / s->dy;
      fsk::k_gradient64<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->p64[s->pidx], s->u[1 - c],
                                                                      s->v[1 - c], s->nx, s->ny, idx, idy, s->red);
      LAUNCHCHK();
      s->cur = 1 - c;
      break;
    }
    default: return fail(FS_E_INVALID_ARG, "unknown kernel %d", kernel);
  }
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  prof_collect(s);
  return FS_OK;
}

int fs_profile_enable(void* h, int kind) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (kind < -1 || kind > 3) return fail(FS_E_INVALID_ARG, "unknown profile kind %d", kind);
  if ((rc = set_device(s))) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  prof_collect(s);
  s->prof.kind = kind;
  s->prof.total_ms = 0;
  s->prof.launches = 0;
  return FS_OK;
}

int fs_profile_read(void* h, double* total_ms, int64_t* launches) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  prof_collect(s);
  if (total_ms) *total_ms = s->prof.total_ms;
  if (launches) *launches = s->prof.launches;
  return FS_OK;
}

}  // extern "C"
