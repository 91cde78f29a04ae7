// fs_solver.hip -- host runtime and C ABI of the MI355X stable-fluids solver.
//
// One solver owns one HIP stream (default priority) and all of its device
// memory; fs_step only enqueues. The step follows src/gpu/solver.js:518-591
// (advect -> diffuse -> force/drag -> project) with the CPU forcing semantics
// of src/cpu/solver.js:353-358 (force, then drag). The per-step coefficients
// (dt, dt/dx, nu dt/dx^2, drag factor) live on the device, so a CFL time step
// is computed and consumed there without a host round trip.
//
// Projection modes
//   jacobi, red-black   FP32 pressure, fixed sweep count, p zeroed per step
//                       (src/gpu/solver.js:658-738).
//   multigrid           iterative refinement: r = div u* - L p in FP64, one
//                       V-cycle on the FP32 correction e (L e = r), p += e in
//                       FP64. A V-cycle is affine in its start value, so k
//                       refinement iterations equal the reference's k V-cycles
//                       on p (src/gpu/solver.js:686-703) up to rounding.
//                       With a tolerance, a predicted number of V-cycles is
//                       enqueued, the per-block residual maxima of every
//                       iteration are read back once, and further V-cycles are
//                       added only if the last residual is still above the gate.
//                       Extra V-cycles only lower the residual, so an
//                       over-prediction costs time, never accuracy.
//
// Also here: the learned pressure guess (fs_guess.hpp) and its trainer
// (fs_train.hpp, fs_trainer_*), batched walled cavities (fs_batch.hpp,
// fs_batch_*) and the on-device energy spectrum (fs_spectrum.hpp).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/fluidsim.h"
#include "fs_kernels.hpp"
#include "fs_mg_fused.hpp"
#include "fs_guess.hpp"
#include "fs_train.hpp"
#include "fs_batch.hpp"
#include "fs_spectrum.hpp"

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

#define HIPCHK(expr)                                                                                  \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess)                                                                             \
      return fail(e_ == hipErrorOutOfMemory ? FS_E_OOM : FS_E_DEVICE, "%s failed: %s (%s:%d)", #expr, \
                  hipGetErrorString(e_), __FILE__, __LINE__);                                         \
  } while (0)

#define LAUNCHCHK()                                                                                  \
  do {                                                                                               \
    hipError_t e_ = hipGetLastError();                                                               \
    if (e_ != hipSuccess)                                                                            \
      return fail(FS_E_DEVICE, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, \
                  __LINE__);                                                                         \
  } while (0)

struct Level {
  int n = 0;
  double dx = 0;
  float* x = nullptr;    // correction (level 0: e); fused path: after pre-smoothing
  float* rhs = nullptr;  // right-hand side (level 0: r)
  float* xb = nullptr;   // fused path: correction after post-smoothing (level 0: e), FP32 under every storage
  int tile[2] = {0, 0};  // fused path: tile shape (TileKind) of the down and up kernels
  float tune_us[2][7] = {};  // autotuner timings per shape (0: not timed)
};

// Kernel classes timed by fs_profile_*: each profiled launch goes through
// hipExtLaunchKernelGGL with a start/stop event pair, which time the dispatch
// itself (not the gap before it), and carries its algorithmic byte count.
enum {
  PK_ADVECT_SL = 0, PK_ADVECT_MC, PK_DIFFUSE, PK_DOWN0, PK_DOWN, PK_COARSE, PK_UP, PK_GRADIENT, PK_RESIDUAL,
  PK_RB, PK_GUESS, PK_ADVECT_FUSED, PK_FORCING, PK_UP_COARSE, PK_COUNT
};

struct Profile {
  unsigned mask = 0;
  struct Pending {
    int kind;
    hipEvent_t a, b;
    double bytes;
  };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> pool;
  double ms[PK_COUNT] = {};
  int64_t launches[PK_COUNT] = {};
  double bytes[PK_COUNT] = {};
};

constexpr int DT_RING = 4096;  // steps whose dt may be pending a host read
constexpr int REC_RING = 64;   // host-visible step records of gated steps
constexpr int MAX_AHEAD = 4;   // gated steps in flight before the host waits for the oldest
constexpr int GATE_SPARE = 1;  // V-cycles enqueued beyond the prediction

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
  int p32idx = 0;            // p32 buffer holding the last flat-solver pressure
  double* p64[2] = {nullptr, nullptr};
  int pidx = 0;              // p64 buffer holding the last projection's pressure
  bool have_p = false;       // a previous pressure exists (warm start)
  float4* dye[2] = {nullptr, nullptr};  // rgba dye at the cell centres (allocated on first use)
  int dye_cur = 0;
  // render(): the rgba half-float image (nx*ny) and the two half-resolution bloom
  // buffers ((nx/2)*(ny/2)), allocated on the first render; bloom_cur holds the last
  // bloom pass's output
  fsk::Half4* out_img = nullptr;
  fsk::Half4* bloom[2] = {nullptr, nullptr};
  int bloom_cur = 0;
  std::vector<fsk::Splat> splats;         // splats of the step being enqueued
  double* p_hist = nullptr;  // FS_GUESS_EXTRAPOLATED: the pressure of the step before the last
  bool have_hist = false;    // p_hist holds it
  std::vector<Level> lv;     // multigrid hierarchy; lv[0].x = e, lv[0].rhs = r
  bool fused = false;        // LDS-tiled V-cycle in use
  bool bf16 = false;         // multigrid correction and right-hand sides stored as BF16
  float* pinv = nullptr;     // coarsest-level pseudo-inverse (null: run the sweeps)
  float* ex_q = nullptr;     // exact coarse solve: the operand table (fsk::EX_TAB 16-byte units)
  bool spare = false;        // launches being enqueued belong to a gated body's spare V-cycle
  bool fold_exact = true;    // the exact solve inside the up launch above it (fold_coarse())
  int coarse_start = 0;      // first level solved by the one-workgroup coarse kernel
  // residual partials: rpart[it * rp_cols + block] = block max |r| of iteration it
  double* rpart = nullptr;
  double* rpart_host = nullptr;  // pinned mirror
  hipEvent_t rb_event = nullptr; // the residual rows of a tolerance-driven projection are on the host
  bool grad_pending = false;     // the gradient of the final pressure is already enqueued
  int rp_rows = 0, rp_cols = 0;
  int resid_row_pending = -1;    // fixed-count projection whose residual is not yet read
  // |u|, |v| block maxima of the last gradient pass, for the CFL step
  double* mpart = nullptr;
  int mp_count = 0;
  bool have_maxabs = false;
  fsk::StepCoef* coef = nullptr;  // device step coefficients
  double* dt_ring = nullptr;      // pinned: dt of steps not yet folded into t
  int dt_pending = 0;
  double* diag_part = nullptr;
  double* diag_out = nullptr;
  double* host = nullptr;         // pinned staging for small reads
  // forcing
  bool forcing_on = false;
  fs_forcing forcing{};
  // the mode list of setForcing(modes, {amplitude, drag, phaseDrift}) (src/gpu/solver.js:
  // 370-414) when forcing_on and mode_list; otherwise the seeded band forcing
  bool mode_list = false;
  std::vector<double> mkx, mky, mph;
  double m_amp = 0, m_drag = 0, m_drift = 0;
  // learned initial guess: weights (BF16 [cout][tap][cin] for the MFMA layers) and activations
  struct GuessNet {
    bool loaded = false;
    float* w0 = nullptr;            // first layer FP32 [16][9][2]
    float* b[15] = {};              // biases (FP32)
    uint16_t* w[15] = {};           // 3x3 layers 1..13 (BF16)
    float* wout = nullptr;          // output 1x1 FP32 [16]
    uint16_t* act[14] = {};         // activations, NHWC BF16
    double* scales = nullptr;       // max|r|, max|p_prev|
    double* ppart = nullptr;        // per-block max|p_prev|
    int nblk = 0;
    int m = 0;                      // resolution the network runs at (n, or n / 2)
    // at half resolution: the averaged inputs, the coarse increment, per-block maxima
    float* rc = nullptr;
    double* pc = nullptr;
    double* dc = nullptr;
    double* cpart = nullptr;
  } guess_net;
  // energy spectrum work space (allocated on first use)
  fsk::cplx* spec_Z = nullptr;
  double* spec_part = nullptr;
  double* spec_E = nullptr;
  // separable forcing tables: MAX_MODES x nx and MAX_MODES x ny (allocated on first use)
  float4 *ftab_col = nullptr, *ftab_row = nullptr;
  // immersed disk: state for the next step, per-block impulse partials, FP64 total
  bool solid_on = false;
  fs_disk solid{};
  double* solid_part = nullptr;
  double* solid_imp = nullptr;
  // stats
  double t = 0;
  int64_t steps = 0;
  int last_vcycles = 0;
  int64_t total_vcycles = 0;
  double last_residual = 0;
  double last_dt = 0;
  int last_pred = 2;  // predicted V-cycles for the next projection
  Profile prof;
  // Device-gated projection (tolerance-driven multigrid on the fused V-cycle): each
  // step enqueues the predicted V-cycles plus a spare, every launch gated by the
  // convergence flags its residual pass leaves, so the host never waits inside a
  // step; the gradient pass records what happened in a host-visible ring, read when
  // the step's event has completed.
  bool gated = true;          // use it (after the first projection has set the prediction)
  bool pred_known = false;    // last_pred comes from a projection of the current flow
  int* gate_flag = nullptr;   // device: one flag per refinement iteration
  int gate_len = 0;
  fsk::Gate gate_r = fsk::gate_open(), gate_v = fsk::gate_open();  // gates of the launches being enqueued
  fsk::ModeSet* modes_dev = nullptr;  // the step's forcing modes (device), written by the step head
  fsk::StepRecord* ring = nullptr;    // host-visible step records, slot = step % RING
  struct Inflight {
    hipEvent_t ev;
    int64_t step;
  };
  std::vector<Inflight> inflight;     // gated steps whose record has not been read
  std::vector<hipEvent_t> ev_pool;
  int64_t gate_misses = 0;            // gated steps whose last residual was above the tolerance
  int64_t gated_steps = 0;
  // captured step bodies, keyed by what changes the launches they hold
  struct StepGraph {
    int k, a0, zero_p, cur, keep, hist, dye;
    int cur_after;  // velocity buffer after the body
    int dye_after;
    hipGraphExec_t exec;
  };
  bool use_graph = true;
  std::vector<StepGraph> graphs;
  int64_t graph_captures = 0, graph_launches = 0;
};

namespace {

inline dim3 grid2(int nx, int ny) {
  return dim3((unsigned)((nx + fsk::BX - 1) / fsk::BX), (unsigned)((ny + fsk::BY - 1) / fsk::BY));
}
inline dim3 grid_pairs(int nx, int ny) {
  int half = (nx + 1) / 2;
  return dim3((unsigned)((half + fsk::BX - 1) / fsk::BX), (unsigned)((ny + fsk::BY - 1) / fsk::BY));
}
inline int grid_blocks(int nx, int ny) {
  dim3 g = grid2(nx, ny);
  return (int)(g.x * g.y);
}
const dim3 BLK(fsk::BX, fsk::BY);

template <typename T>
int dalloc(T** p, size_t count) {
  size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 1;
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(e == hipErrorOutOfMemory ? FS_E_OOM : FS_E_DEVICE, "hipMalloc(%zu bytes) failed: %s", bytes,
                hipGetErrorString(e));
  }
  return FS_OK;
}

template <typename T>
int halloc(T** p, size_t count) {
  size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 1;
  if (hipHostMalloc((void**)p, bytes, hipHostMallocDefault) != hipSuccess) {
    *p = nullptr;
    return fail(FS_E_OOM, "hipHostMalloc(%zu bytes) failed", bytes);
  }
  return FS_OK;
}

// Host memory the device writes during a kernel and the host reads once the kernel's
// completion has been observed (step records): coherent, so no cache holds the writes.
template <typename T>
int halloc_coherent(T** p, size_t count) {
  size_t bytes = count * sizeof(T) > 0 ? count * sizeof(T) : 1;
  if (hipHostMalloc((void**)p, bytes, hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
    *p = nullptr;
    return fail(FS_E_OOM, "hipHostMalloc(%zu bytes, coherent) failed", bytes);
  }
  std::memset((void*)*p, 0xff, bytes);  // no record claims a step yet
  return FS_OK;
}

void dfree(void* p) {
  if (p) (void)hipFree(p);
}
void hfree(void* p) {
  if (p) (void)hipHostFree(p);
}

hipEvent_t take_event(fs_solver_t* s) {
  if (!s->prof.pool.empty()) {
    hipEvent_t e = s->prof.pool.back();
    s->prof.pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

// Launch on the solver's stream; when `kind` is being profiled the dispatch is
// timed by events recorded at its own start and end.
template <typename K, typename... Args>
int launch(fs_solver_t* s, int kind, double bytes, K kern, dim3 g, dim3 b, Args... args) {
  if (kind >= 0 && ((s->prof.mask >> kind) & 1u)) {
    hipEvent_t e0 = take_event(s), e1 = take_event(s);
    hipExtLaunchKernelGGL(kern, g, b, 0, s->stream, e0, e1, 0, args...);
    s->prof.pending.push_back({kind, e0, e1, bytes});
  } else {
    kern<<<g, b, 0, s->stream>>>(args...);
  }
  hipError_t e_ = hipGetLastError();
  if (e_ != hipSuccess) return fail(FS_E_DEVICE, "kernel launch failed: %s", hipGetErrorString(e_));
  return FS_OK;
}

// Everything the host learns lazily after a stream synchronisation.
void after_sync(fs_solver_t* s) {
  // launches still running (after a wait on an event rather than the stream) stay
  // pending until a later synchronisation
  size_t keep = 0;
  for (auto& pr : s->prof.pending) {
    if (hipEventQuery(pr.b) != hipSuccess) {
      s->prof.pending[keep++] = pr;
      continue;
    }
    float ms = 0;
    if (hipEventElapsedTime(&ms, pr.a, pr.b) == hipSuccess) {
      s->prof.ms[pr.kind] += ms;
      s->prof.launches[pr.kind]++;
      s->prof.bytes[pr.kind] += pr.bytes;
    }
    s->prof.pool.push_back(pr.a);
    s->prof.pool.push_back(pr.b);
  }
  s->prof.pending.resize(keep);
  for (int k = 0; k < s->dt_pending; k++) {
    s->t += s->dt_ring[k];
    s->last_dt = s->dt_ring[k];
  }
  s->dt_pending = 0;
  if (s->resid_row_pending >= 0) {
    // the fixed-count residual row was copied before the synchronisation
    double m = 0;
    const double* row = s->rpart_host + (size_t)s->resid_row_pending * s->rp_cols;
    for (int b = 0; b < s->rp_cols; b++) m = std::max(m, row[b]);
    s->last_residual = m;
    s->resid_row_pending = -1;
  }
}

int consume_records(fs_solver_t* s, bool wait);

// Wait for the stream; then every timing, dt and step record is on the host.
int sync_stream(fs_solver_t* s) {
  HIPCHK(hipStreamSynchronize(s->stream));
  after_sync(s);
  return consume_records(s, true);
}

int check_solver(void* h) {
  if (!h) return fail(FS_E_INVALID_ARG, "solver handle is NULL");
  return FS_OK;
}

int set_device(fs_solver_t* s) {
  HIPCHK(hipSetDevice(s->dev));
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

double drag_alpha(const fs_solver_t* s) {
  return !s->forcing_on ? 0.0 : (s->mode_list ? s->m_drag : s->forcing.drag);
}

// ------------------------------------------------------------- step pieces

double js_round(double x) {
  double r = std::floor(x);
  if (x - r >= 0.5) r += 1.0;
  return r;
}

// The band-forcing modes of step `step` (src/cpu/solver.js:181-199) from the exact
// uint32 LCG; the amplitude factor dt and the drag factor come from the device
// coefficients.
void make_modes(const fs_solver_t* s, int64_t step, fsk::ModeSet& ms) {
  if (s->mode_list) {
    // FORCING (src/gpu/shaders.js:532-566): every mode in the list counts in sqrt(count);
    // a zero wavevector adds nothing
    ms = fsk::ModeSet{};
    const int M = (int)s->mkx.size();
    const double scale = s->m_amp / std::sqrt((double)M);
    int k = 0;
    for (int m = 0; m < M; m++) {
      const double kx = s->mkx[m], ky = s->mky[m], kk = std::sqrt(kx * kx + ky * ky);
      if (kk < 1e-6) continue;
      fsk::Mode& md = ms.m[k++];
      md.tx = kx * s->dx / (2 * M_PI);
      md.ty = ky * s->dy / (2 * M_PI);
      md.tph = s->mph[m] / (2 * M_PI);
      md.cu = (float)(scale * (ky / kk));
      md.cv = (float)(scale * (-kx / kk));
    }
    ms.n = k;
    ms.after_drag = 1;
    ms.drift_turns = s->m_drift / (2 * M_PI);
    return;
  }
  const fs_forcing& f = s->forcing;
  uint32_t st = f.seed + (uint32_t)step;
  auto rnd = [&]() {
    st = st * 1664525u + 1013904223u;
    return (double)st / 4294967296.0;
  };
  ms = fsk::ModeSet{};
  const int M = 24;
  double scale = f.amp / std::sqrt((double)M);
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
}

// The first launch of a step: dt (fixed, or from the CFL condition on the device),
// the coefficients and step index, the step's forcing modes, cleared gate flags.
int enqueue_head(fs_solver_t* s, bool use_cfl, double value, int nflags, bool advance_t = true) {
  if (use_cfl && !s->have_maxabs) {
    fsk::k_maxabs<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[s->cur], s->v[s->cur], s->nx, s->ny, s->mpart);
    LAUNCHCHK();
  }
  fsk::ModeSet ms{};
  if (s->forcing_on) make_modes(s, s->steps, ms);
  fsk::k_step_head<<<1, 1024, 0, s->stream>>>(s->mpart, s->mp_count, use_cfl ? 1 : 0, value, s->dx, s->dy,
                                              s->cfg.nu, drag_alpha(s), 1e30, s->steps, advance_t ? 1 : 0, s->coef,
                                              s->gate_flag,
                                              std::min(nflags, s->gate_len), ms, s->modes_dev,
                                              s->forcing_on ? 1 : 0);
  LAUNCHCHK();
  return FS_OK;
}

int enqueue_advection(fs_solver_t* s) {
  int c = s->cur, rc;
  const double N = (double)s->n;
  if (s->cfg.advection == FS_ADV_MACCORMACK) {
    if ((rc = launch(s, PK_ADVECT_SL, 16 * N, fsk::k_advect_sl, grid2(s->nx, s->ny), BLK, (const float*)s->u[c],
                     (const float*)s->v[c], s->fu, s->fv, s->nx, s->ny, (const fsk::StepCoef*)s->coef, 1.0f)))
      return rc;
    if ((rc = launch(s, PK_ADVECT_MC, 24 * N, fsk::k_advect_mc, grid2(s->nx, s->ny), BLK, (const float*)s->u[c],
                     (const float*)s->v[c], (const float*)s->fu, (const float*)s->fv, s->u[1 - c], s->v[1 - c], s->nx,
                     s->ny, (const fsk::StepCoef*)s->coef)))
      return rc;
  } else {
    if ((rc = launch(s, PK_ADVECT_SL, 16 * N, fsk::k_advect_sl, grid2(s->nx, s->ny), BLK, (const float*)s->u[c],
                     (const float*)s->v[c], s->u[1 - c], s->v[1 - c], s->nx, s->ny, (const fsk::StepCoef*)s->coef,
                     1.0f)))
      return rc;
  }
  s->cur = 1 - c;
  return FS_OK;
}

// MacCormack advection and explicit diffusion in one tiled launch, the same bits as
// k_advect_sl + k_advect_mc + k_diffuse; 16 B per cell compulsory (u, v in and out).
int enqueue_advect_diffuse_fused(fs_solver_t* s) {
  const int c = s->cur;
  const int tx = (s->nx + fsk::AT - 1) / fsk::AT, ty = (s->ny + fsk::AT - 1) / fsk::AT, nb = tx * ty;
  int rc = launch(s, PK_ADVECT_FUSED, 16.0 * s->n, fsk::k_advect_mc_diffuse, dim3((unsigned)nb), dim3(fsk::ANT),
                  (const float*)s->u[c], (const float*)s->v[c], s->u[1 - c], s->v[1 - c], s->nx, s->ny, tx, nb,
                  (const fsk::StepCoef*)s->coef, s->cfg.nu > 0 ? 1 : 0);
  if (rc) return rc;
  s->cur = 1 - c;
  return FS_OK;
}

int enqueue_diffusion(fs_solver_t* s) {
  if (!(s->cfg.nu > 0)) return FS_OK;
  int c = s->cur;
  int rc = launch(s, PK_DIFFUSE, 16.0 * s->n, fsk::k_diffuse, grid2(s->nx, s->ny), BLK, (const float*)s->u[c],
                  (const float*)s->v[c], s->u[1 - c], s->v[1 - c], s->nx, s->ny, (const fsk::StepCoef*)s->coef);
  if (rc) return rc;
  s->cur = 1 - c;
  return FS_OK;
}

// Band forcing then drag of src/cpu/solver.js:200-237 with the modes the step head
// copied to the device: per-step cos/sin tables, then the separable update.
int enqueue_forcing(fs_solver_t* s) {
  if (!s->forcing_on) return FS_OK;
  int c = s->cur, rc;
  // sized for every mode: the count of the step is on the device (a mode whose rounded
  // wavenumber is zero is dropped); with no mode in the band the pass below still drags
  const unsigned tb = (unsigned)((fsk::MAX_MODES * (s->nx + s->ny) + 255) / 256);
  if ((rc = launch(s, PK_FORCING, 16.0 * fsk::MAX_MODES * (s->nx + s->ny), fsk::k_forcing_tables, dim3(tb), dim3(256),
                   (const fsk::ModeSet*)s->modes_dev, s->nx, s->ny, s->ftab_col, s->ftab_row,
                   (const fsk::StepCoef*)s->coef)))
    return rc;
  if (s->nx % fsk::FMX == 0 && s->ny % fsk::FMX == 0)  // the field as a product on the matrix cores
    return launch(s, PK_FORCING, 16.0 * s->n, fsk::k_forcing_mma,
                  dim3((unsigned)(s->nx / fsk::FMX), (unsigned)(s->ny / fsk::FMX)), dim3(256), s->u[c], s->v[c],
                  s->nx, s->ny, (const fsk::ModeSet*)s->modes_dev, (const float4*)s->ftab_col,
                  (const float4*)s->ftab_row, (const fsk::StepCoef*)s->coef);
  const dim3 g((unsigned)((s->nx + fsk::FX - 1) / fsk::FX), (unsigned)((s->ny + fsk::FY - 1) / fsk::FY));
  return launch(s, PK_FORCING, 16.0 * s->n, fsk::k_forcing_sep, g, dim3(256), s->u[c], s->v[c], s->nx, s->ny,
                (const fsk::ModeSet*)s->modes_dev, (const float4*)s->ftab_col, (const float4*)s->ftab_row,
                (const fsk::StepCoef*)s->coef);
}

// The step's splats (after the forcing, src/gpu/solver.js:546-554), up to MAX_SPLATS per launch.
int enqueue_splats(fs_solver_t* s) {
  for (size_t at = 0; at < s->splats.size(); at += fsk::MAX_SPLATS) {
    fsk::SplatSet ss{};
    ss.n = (int)std::min((size_t)fsk::MAX_SPLATS, s->splats.size() - at);
    for (int m = 0; m < ss.n; m++) ss.s[m] = s->splats[at + m];
    fsk::k_splat<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[s->cur], s->v[s->cur],
                                                             s->dye[0] ? s->dye[s->dye_cur] : nullptr, s->nx, s->ny,
                                                             (float)s->dx, (float)s->dy, ss);
    LAUNCHCHK();
  }
  return FS_OK;
}

// Vorticity confinement after the splats and before the disk and the projection
// (src/gpu/solver.js:558-568), when its strength is positive.
int enqueue_confinement(fs_solver_t* s) {
  if (!(s->cfg.confinement > 0)) return FS_OK;
  const int c = s->cur;
  fsk::k_confine<<<fsk::vt_grid(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->u[1 - c], s->v[1 - c], s->nx,
                                                             s->ny, (float)s->dx, (float)s->dy,
                                                             (float)s->cfg.confinement, (const fsk::StepCoef*)s->coef);
  LAUNCHCHK();
  s->cur = 1 - c;
  return FS_OK;
}

// The dye through the projected velocity (the step's last pass, src/gpu/solver.js:575-584).
int enqueue_dye(fs_solver_t* s) {
  if (!s->dye[0]) return FS_OK;
  const int c = s->dye_cur;
  fsk::k_advect_dye<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(
      s->u[s->cur], s->v[s->cur], s->dye[c], s->dye[1 - c], s->nx, s->ny, (const fsk::StepCoef*)s->coef,
      (float)s->cfg.dye_fade);
  LAUNCHCHK();
  s->dye_cur = 1 - c;
  return FS_OK;
}

// Direct-forcing coupling of the disk, after forcing and before the projection
// (src/cpu/solver.js:373-383; src/gpu/solver.js:570-573).
int enqueue_solid(fs_solver_t* s) {
  if (!s->solid_on) return FS_OK;
  const fs_disk& d = s->solid;
  fsk::Disk k{};
  k.x = d.x;
  k.y = d.y;
  k.r = d.r;
  k.vx = d.vx;
  k.vy = d.vy;
  k.omega = d.omega;
  k.width = s->cfg.solid_mask_width * s->dx;
  k.dA = s->dx * s->dy;
  k.lx = s->cfg.lx;
  k.ly = s->cfg.ly;
  k.hx = s->dx;
  k.hy = s->dy;
  k.passes = d.passes;
  int c = s->cur;
  dim3 g = grid2(s->nx, s->ny);
  fsk::k_solid<<<g, BLK, 0, s->stream>>>(s->u[c], s->v[c], s->nx, s->ny, k, s->solid_part);
  LAUNCHCHK();
  fsk::k_solid_final<<<1, 256, 0, s->stream>>>(s->solid_part, (int)(g.x * g.y), s->solid_imp);
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

// Jacobi / red-black with FP32 pressure; the result index comes back in out_idx.
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
        rc = launch(s, PK_RB, 6.0 * s->n, fsk::k_rb_colour<float>, grid_pairs(s->nx, s->ny), BLK, s->p32[0],
                    (const float*)s->div, s->nx, s->ny, h2, 1.0f, colour, 0);
        if (rc) return rc;
      }
    }
  }
  *out_idx = a;
  return FS_OK;
}

// ---------------------------------------------------- per-operator V-cycle

// Table of levels from..end-1 for the one-workgroup coarse kernel.
fsk::LevelTable level_table(const fs_solver_t* s, int from, int end) {
  fsk::LevelTable tb{};
  int o = 0;
  tb.levels = end - from;
  tb.tail = tb.levels;
  for (int l = from; l < end; l++) {
    int k = l - from;
    double dx = s->lv[l].dx;
    tb.h2[k] = (float)(dx * dx);
    tb.ih2[k] = (float)(1.0 / (dx * dx));
    tb.n[k] = s->lv[l].n;
    tb.mn[k] = fsk::qdiv_magic(s->lv[l].n);
    tb.mh[k] = fsk::qdiv_magic(std::max(1, s->lv[l].n / 2));
    tb.off[k] = o;
    o += s->lv[l].n * s->lv[l].n;
    if (tb.tail == tb.levels && s->lv[l].n <= fsk::REG_N) tb.tail = k;
  }
  // 16 x 16 and 32 x 32 levels just above an 8 x 8 register tail run as register blocks
  tb.wave = tb.tail;
  if (tb.tail < tb.levels && tb.n[tb.tail] == 8 && tb.tail >= 1 && tb.n[tb.tail - 1] == 16) {
    tb.wave = tb.tail - 1;
    if (tb.tail >= 2 && tb.n[tb.tail - 2] == 32) tb.wave = tb.tail - 2;
  }
  return tb;
}

// `sweeps` red-black sweeps on level l; x_zero marks a logically-zero start.
int mg_smooth(fs_solver_t* s, int l, int sweeps, bool& x_zero) {
  Level& L = s->lv[l];
  float h2 = (float)(L.dx * L.dx);
  for (int k = 0; k < sweeps; k++) {
    for (int colour = 0; colour < 2; colour++) {
      int rc = launch(s, PK_RB, 6.0 * L.n * L.n, fsk::k_mg_rb<float>, grid_pairs(L.n, L.n), BLK, L.x,
                      (const float*)L.rhs, L.n, h2, colour, x_zero ? 1 : 0);
      if (rc) return rc;
      // after the red pass from zero, black cells are still stale but the black
      // pass never reads its own value, and red cells are now valid
      x_zero = false;
    }
  }
  return FS_OK;
}

// One V-cycle on the correction (lv[0].x = e, lv[0].rhs = r), one launch per operator.
int mg_vcycle(fs_solver_t* s) {
  int Lc = (int)s->lv.size();
  std::vector<char> zero(Lc, 1);  // every level's correction starts from zero
  for (int l = 0; l < Lc - 1; l++) {
    bool z = zero[l];
    int rc = mg_smooth(s, l, s->cfg.mg_pre, z);
    if (rc) return rc;
    zero[l] = z;
    Level& F = s->lv[l];
    Level& Cl = s->lv[l + 1];
    fsk::k_mg_restrict<float><<<grid2(Cl.n, Cl.n), BLK, 0, s->stream>>>(F.x, F.rhs, F.n, (float)(1.0 / (F.dx * F.dx)),
                                                                         Cl.rhs, zero[l] ? 1 : 0);
    LAUNCHCHK();
  }
  {
    Level& C = s->lv[Lc - 1];
    if (s->ex_q) {
      int rc = launch(s, PK_COARSE, 8.0 * C.n * C.n + 16.0 * fsk::EX_TAB, fsk::k_mg_coarse_exact<float>,
                      dim3(1), dim3(fsk::EXT), C.x, (const float*)C.rhs, C.n, (const long*)nullptr,
                      (const uint4*)s->ex_q, fsk::gate_open());
      if (rc) return rc;
    } else if (C.n <= fsk::COARSE_N) {
      int rc = launch(s, PK_COARSE, 8.0 * C.n * C.n, fsk::k_mg_coarse_ml<float>, dim3(1), dim3(fsk::CTH), C.x,
                      (const float*)C.rhs, level_table(s, Lc - 1, Lc), s->cfg.mg_pre, s->cfg.mg_post,
                      s->cfg.mg_coarse, (const float*)s->pinv, fsk::gate_open());
      if (rc) return rc;
    } else {
      bool z = true;
      int rc = mg_smooth(s, Lc - 1, s->cfg.mg_coarse, z);
      if (rc) return rc;
      if (z) {  // zero sweeps: the coarse correction is zero
        rc = fill_f32(s, C.x, (size_t)C.n * C.n, 0.f);
        if (rc) return rc;
      }
    }
  }
  for (int l = Lc - 2; l >= 0; l--) {
    Level& F = s->lv[l];
    Level& Cl = s->lv[l + 1];
    fsk::k_mg_prolong<float><<<grid2(F.n, F.n), BLK, 0, s->stream>>>(Cl.x, F.x, F.n, zero[l] ? 1 : 0);
    LAUNCHCHK();
    bool z = false;
    int rc = mg_smooth(s, l, s->cfg.mg_post, z);
    if (rc) return rc;
  }
  return FS_OK;
}

// ----------------------------------------------------------- fused V-cycle

// Tile shapes of the grid-wide levels. The default is the largest shape that still
// yields a tile per CU; the autotuner replaces it with the fastest measured one.
// Every shape computes the same bits (each cell sees the same operands in the same
// order), so the choice affects time only.
// Shapes: 64 x 32 (512 threads), 32 x 32 (256), 16 x 16 (128), 8 x 8 (64) (the first
// TILE_KINDS), 64 x 64 tiles of 1024 threads, and 32 x 32 / 16 x 16 tiles with four times /
// twice the threads.
enum TileKind { TILE64 = 0, TILE32, TILE16, TILE8, TILE_KINDS, TILE64SQ = TILE_KINDS, TILE32W, TILE16W,
                GRID_TILE_KINDS };
static const int kTileW[GRID_TILE_KINDS] = {64, 32, 16, 8, 64, 32, 16};
static const int kTileH[GRID_TILE_KINDS] = {32, 32, 16, 8, 64, 32, 16};
static const char* kTileName[GRID_TILE_KINDS] = {"64x32", "32x32", "16x16", "8x8", "64x64", "32x32/1024",
                                                 "16x16/256"};

inline int tile_count(int k, int n) { return ((n + kTileW[k] - 1) / kTileW[k]) * ((n + kTileH[k] - 1) / kTileH[k]); }

inline int tile_kind(int n) {
  for (int k = TILE64; k < TILE8; k++)
    if (tile_count(k, n) >= 256) return k;
  return TILE8;
}

template <class TS>
inline dim3 tiles_grid_t(int n, int* tiles_x, int* nblocks) {
  int tx = (n + TS::TX - 1) / TS::TX, ty = (n + TS::TY - 1) / TS::TY;
  *tiles_x = tx;
  *nblocks = tx * ty;
  return dim3((unsigned)(tx * ty));
}

#define DISPATCH_TILE(k, fn, SW, ...)                                                        \
  ((k) == TILE64     ? fn<SW, fsk::Tile64>(__VA_ARGS__)                                       \
   : (k) == TILE32   ? fn<SW, fsk::Tile32>(__VA_ARGS__)                                       \
   : (k) == TILE16   ? fn<SW, fsk::Tile16>(__VA_ARGS__)                                       \
   : (k) == TILE64SQ ? fn<SW, fsk::Tile64sq>(__VA_ARGS__)                                     \
   : (k) == TILE32W  ? fn<SW, fsk::Tile32w>(__VA_ARGS__)                                      \
   : (k) == TILE16W  ? fn<SW, fsk::Tile16w>(__VA_ARGS__)                                      \
                     : fn<SW, fsk::Tile8>(__VA_ARGS__))

template <typename S>
inline S* as(float* p) {
  return reinterpret_cast<S*>(p);
}
template <typename S>
inline const S* as(const float* p) {
  return reinterpret_cast<const S*>(p);
}

template <int PRE>
int launch_down0(fs_solver_t* s, int it, int pin, bool zero_p, bool residual_only);
template <int PRE, class TS, typename S>
int launch_down0_s(fs_solver_t* s, int it, int pin, bool zero_p, bool residual_only) {
  Level& L0 = s->lv[0];
  int tx, nb;
  dim3 g = tiles_grid_t<TS>(L0.n, &tx, &nb);
  double idx = 1.0 / s->dx;
  int c = s->cur;
  const double es = sizeof(S);
  // u, v (8) + p (8) + previous correction (FP32) read; p (8) + r written; then x and the coarse rhs (1/4 cell)
  double bpc = 8 + (zero_p ? 0 : 8) + (it > 0 ? 4 : 0) + 8 + es + (residual_only ? 0 : es + 0.25 * es);
  // 16-byte accesses and fixed thread maps when tiles are whole and quads stay aligned
  if constexpr ((2 * PRE) % 4 == 0 && TS::TX % 4 == 0) {
    const int n = L0.n;
    if (n % 4 == 0 && n % TS::TX == 0 && n % TS::TY == 0 && n <= 32768)
      return launch(s, PK_DOWN0, bpc * L0.n * L0.n, fsk::k_mgt_down0q<PRE, TS, S>, g, dim3(TS::NT),
                    (const float*)s->u[c], (const float*)s->v[c], (const double*)(zero_p ? nullptr : s->p64[pin]),
                    (const float*)(it > 0 ? L0.xb : nullptr), s->p64[1 - pin], as<S>(L0.rhs), as<S>(L0.x),
                    as<S>(s->lv[1].rhs), L0.n, idx, idx * idx, (float)(L0.dx * L0.dx),
                    (float)(1.0 / (L0.dx * L0.dx)), tx, nb, s->rpart + (size_t)it * s->rp_cols,
                    residual_only ? 1 : 0, s->gate_r);
  }
  return launch(s, PK_DOWN0, bpc * L0.n * L0.n, fsk::k_mgt_down0<PRE, TS, S>, g, dim3(TS::NT), (const float*)s->u[c],
                (const float*)s->v[c], (const double*)(zero_p ? nullptr : s->p64[pin]),
                (const float*)(it > 0 ? L0.xb : nullptr), s->p64[1 - pin], as<S>(L0.rhs), as<S>(L0.x),
                as<S>(s->lv[1].rhs), L0.n, idx, idx * idx, (float)(L0.dx * L0.dx), (float)(1.0 / (L0.dx * L0.dx)), tx,
                nb, s->rpart + (size_t)it * s->rp_cols, residual_only ? 1 : 0, s->gate_r);
}
template <int PRE, class TS>
int launch_down0_t(fs_solver_t* s, int it, int pin, bool zero_p, bool residual_only) {
  return s->bf16 ? launch_down0_s<PRE, TS, fsk::bf16>(s, it, pin, zero_p, residual_only)
                 : launch_down0_s<PRE, TS, float>(s, it, pin, zero_p, residual_only);
}
template <int PRE>
int launch_down0(fs_solver_t* s, int it, int pin, bool zero_p, bool residual_only) {
  return DISPATCH_TILE(s->lv[0].tile[0], launch_down0_t, PRE, s, it, pin, zero_p, residual_only);
}

template <int PRE, class TS, typename S>
int launch_down_s(fs_solver_t* s, int l) {
  Level& F = s->lv[l];
  int tx, nb;
  dim3 g = tiles_grid_t<TS>(F.n, &tx, &nb);
  // rhs read; x and the coarse rhs (1 per 4 cells) written
  // quad accesses when the halo 2 PRE is a multiple of 4 and the level is whole tiles
  auto kern = [&] {
    if constexpr ((2 * PRE) % 4 == 0 && TS::TX % 4 == 0)
      if (F.n % 4 == 0 && F.n % TS::TX == 0 && F.n % TS::TY == 0) return fsk::k_mgt_downq<PRE, TS, S>;
    return fsk::k_mgt_down<PRE, TS, S>;
  }();
  return launch(s, PK_DOWN, 2.25 * sizeof(S) * F.n * F.n, kern, g, dim3(TS::NT), as<S>(F.x),
                as<S>((const float*)F.rhs), F.n, (float)(F.dx * F.dx), (float)(1.0 / (F.dx * F.dx)),
                as<S>(s->lv[l + 1].rhs), tx, nb, s->gate_v);
}
template <int PRE, class TS>
int launch_down_t(fs_solver_t* s, int l) {
  return s->bf16 ? launch_down_s<PRE, TS, fsk::bf16>(s, l) : launch_down_s<PRE, TS, float>(s, l);
}
template <int PRE>
int launch_down(fs_solver_t* s, int l) {
  // a spare V-cycle, gated off unless the prediction fell short, takes the shape with the
  // fewest workgroups (a gated-off launch costs its dispatch, which grows with them)
  const int n = s->lv[l].n;
  const int k = s->spare && n % 64 == 0 ? (int)TILE64SQ : s->lv[l].tile[0];
  return DISPATCH_TILE(k, launch_down_t, PRE, s, l);
}

// Levels l and l + 1 down in one launch (k_mgt_down2), tiles of 8 x 8 cells of level l + 1
// (1024 threads: the launch is latency-bound, and more threads make its passes shorter).
template <int PRE, typename S>
int launch_down2_s(fs_solver_t* s, int l) {
  using TS = fsk::Tile8x1024;
  Level& F = s->lv[l];
  Level& C = s->lv[l + 1];
  int tx, nb;
  dim3 g = tiles_grid_t<TS>(C.n, &tx, &nb);
  // level l: rhs read, x written; level l + 1: x and rhs written; level l + 2 rhs written
  const double bytes = 2.0 * sizeof(S) * F.n * F.n + 2.25 * sizeof(S) * C.n * C.n;
  auto kern = [&] {
    if constexpr ((2 * PRE) % 4 == 0)
      if (C.n % TS::TX == 0 && C.n % TS::TY == 0) return fsk::k_mgt_down2q<PRE, TS, S>;
    return fsk::k_mgt_down2<PRE, TS, S>;
  }();
  return launch(s, PK_DOWN, bytes, kern, g, dim3(TS::NT), as<S>(F.x),
                as<S>((const float*)F.rhs), F.n, (float)(F.dx * F.dx), (float)(1.0 / (F.dx * F.dx)), as<S>(C.x),
                as<S>(C.rhs), (float)(C.dx * C.dx), (float)(1.0 / (C.dx * C.dx)), as<S>(s->lv[l + 2].rhs), tx, nb,
                s->gate_v);
}
template <int PRE>
int launch_down2(fs_solver_t* s, int l) {
  return s->bf16 ? launch_down2_s<PRE, fsk::bf16>(s, l) : launch_down2_s<PRE, float>(s, l);
}

template <int POST, class TS, typename S>
int launch_up_s(fs_solver_t* s, int l) {
  Level& F = s->lv[l];
  int tx, nb;
  dim3 g = tiles_grid_t<TS>(F.n, &tx, &nb);
  // x, rhs (storage precision) and the coarse correction (FP32, 1 per 4 cells) read;
  // the correction written in FP32 (see Level::xb)
  auto kern = [&] {
    if constexpr ((2 * POST) % 4 == 0 && TS::TX % 4 == 0)
      if (F.n % 4 == 0 && F.n % TS::TX == 0 && F.n % TS::TY == 0) return fsk::k_mgt_upq<POST, TS, S, float>;
    return fsk::k_mgt_up<POST, TS, S, float>;
  }();
  return launch(s, PK_UP, (2.0 * sizeof(S) + 5.0) * F.n * F.n, kern, g, dim3(TS::NT),
                as<S>((const float*)F.x), F.xb, as<S>((const float*)F.rhs), (const float*)s->lv[l + 1].xb, F.n,
                (float)(F.dx * F.dx), tx, nb, s->gate_v);
}
// Levels l + 1 and l up in one launch (k_mgt_up2), tiles TS of level l. FOLD: level
// l + 2 is the 64 x 64 exactly solved level, solved inside the launch (fold_coarse).
template <int POST, class TS, typename S, bool FOLD>
int launch_up2_f(fs_solver_t* s, int l) {
  Level& F = s->lv[l];
  Level& C = s->lv[l + 1];
  Level& C2 = s->lv[l + 2];
  int tx, nb;
  dim3 g = tiles_grid_t<TS>(F.n, &tx, &nb);
  // level l + 1: x, rhs, the level-(l + 2) correction (or its rhs and the solve's table)
  // read; level l: x, rhs read, correction written
  const double bytes = (2.0 * sizeof(S) + (FOLD ? 0.0 : 1.0)) * C.n * C.n + (2.0 * sizeof(S) + 4.0) * F.n * F.n +
                       (FOLD ? sizeof(S) * C2.n * C2.n + 16.0 * fsk::EX_TAB : 0.0);
  // quad accesses when the halo 2 POST is a multiple of 4 (levels l and l + 1 are whole
  // multiples of 16 wherever k_mgt_up2 runs)
  auto kern = [] {
    if constexpr ((2 * POST) % 4 == 0) return fsk::k_mgt_up2q<POST, TS, S, float, FOLD>;
    else return fsk::k_mgt_up2<POST, TS, S, float, FOLD>;
  }();
  return launch(s, FOLD ? PK_UP_COARSE : PK_UP, bytes, kern, g, dim3(TS::NT), as<S>((const float*)C.x),
                as<S>((const float*)C.rhs),
                (const float*)C2.xb, as<S>((const float*)F.x), F.xb, as<S>((const float*)F.rhs), F.n,
                (float)(F.dx * F.dx), (float)(C.dx * C.dx), tx, nb, s->gate_v, as<S>((const float*)C2.rhs),
                (const uint4*)s->ex_q);
}
// The exact 64 x 64 solve of the coarse level runs inside the up launch above it (every
// workgroup solves it for its own footprint) instead of as a launch of its own.
bool fold_coarse(const fs_solver_t* s, int l) {
  return s->fold_exact && s->ex_q && l + 2 == s->coarse_start && s->lv[l + 2].n == fsk::EXN;
}
template <int POST, class TS, typename S>
int launch_up2_s(fs_solver_t* s, int l) {
  // the folded solve runs on 4 or 16 wavefronts
  if constexpr (TS::NT == 256 || TS::NT == 1024)
    if (fold_coarse(s, l)) return launch_up2_f<POST, TS, S, true>(s, l);
  return launch_up2_f<POST, TS, S, false>(s, l);
}
// 8 x 8 tiles of 1024 threads for the pair above the coarse solve (levels 256 and 128
// over a 64 coarse level: more workgroups in flight than 16 x 16, short passes), 16 x 16
// of 1024 threads when that pair also solves the coarse level itself, 32 x 32 or (as
// tuned, on a grid of whole 64 x 64 tiles) 64 x 64 for levels 1 and 0
bool up2_64sq_ok(const fs_solver_t* s) { return s->lv[0].n % 64 == 0; }
template <int POST>
int launch_up2(fs_solver_t* s, int l) {
  if (l == 0 && s->lv[0].tile[1] == TILE64SQ && up2_64sq_ok(s))
    return s->bf16 ? launch_up2_s<POST, fsk::Tile64sq, fsk::bf16>(s, l)
                   : launch_up2_s<POST, fsk::Tile64sq, float>(s, l);
  if (l == 0 && s->lv[0].tile[1] == TILE64 && up2_64sq_ok(s) && !fold_coarse(s, 0))
    return s->bf16 ? launch_up2_s<POST, fsk::Tile64, fsk::bf16>(s, l) : launch_up2_s<POST, fsk::Tile64, float>(s, l);
  if (l == 0)
    return s->bf16 ? launch_up2_s<POST, fsk::Tile32, fsk::bf16>(s, l) : launch_up2_s<POST, fsk::Tile32, float>(s, l);
  // with the exact solve folded in, 16 x 16 tiles: a quarter of the workgroups repeat it
  if (fold_coarse(s, l))
    return s->bf16 ? launch_up2_s<POST, fsk::Tile16x1024, fsk::bf16>(s, l)
                   : launch_up2_s<POST, fsk::Tile16x1024, float>(s, l);
  return s->bf16 ? launch_up2_s<POST, fsk::Tile8x1024, fsk::bf16>(s, l)
                 : launch_up2_s<POST, fsk::Tile8x1024, float>(s, l);
}

template <int POST, class TS>
int launch_up_t(fs_solver_t* s, int l) {
  return s->bf16 ? launch_up_s<POST, TS, fsk::bf16>(s, l) : launch_up_s<POST, TS, float>(s, l);
}
template <int POST>
int launch_up(fs_solver_t* s, int l) {
  return DISPATCH_TILE(s->lv[l].tile[1], launch_up_t, POST, s, l);
}

template <typename S>
int launch_coarse_s(fs_solver_t* s, int from) {
  Level& C = s->lv[from];
  // rhs at the storage precision in, the result (the correction passed up) out in FP32
  if (s->ex_q)
    return launch(s, PK_COARSE, (sizeof(S) + 4.0) * C.n * C.n + 16.0 * fsk::EX_TAB,
                  fsk::k_mg_coarse_exact<S>, dim3(1), dim3(fsk::EXT), C.xb, as<S>((const float*)C.rhs), C.n,
                  (const long*)nullptr, (const uint4*)s->ex_q, s->gate_v);
  return launch(s, PK_COARSE, (sizeof(S) + 4.0) * C.n * C.n, fsk::k_mg_coarse_ml<S, float>, dim3(1),
                dim3(fsk::CTH), C.xb, as<S>((const float*)C.rhs), level_table(s, from, (int)s->lv.size()),
                s->cfg.mg_pre, s->cfg.mg_post, s->cfg.mg_coarse, (const float*)s->pinv, s->gate_v);
}
int launch_coarse_ml(fs_solver_t* s, int from) {
  return s->bf16 ? launch_coarse_s<fsk::bf16>(s, from) : launch_coarse_s<float>(s, from);
}

#define DISPATCH_SW(sw, fn, ...) \
  ((sw) == 1 ? fn<1>(__VA_ARGS__) : (sw) == 2 ? fn<2>(__VA_ARGS__) : fn<3>(__VA_ARGS__))

// FP64 residual `it` on the per-operator path (and the fused path when the whole
// hierarchy fits one workgroup): reads p64[pin] (+ e), writes p64[1-pin] and r.
template <typename S>
int enqueue_residual_s(fs_solver_t* s, int it, int pin, bool zero_p, const float* e) {
  int c = s->cur;
  double idx = 1.0 / s->dx, idy = 1.0 / s->dy;
  double bpc = 8 + (zero_p ? 0 : 8) + (e ? 4 : 0) + 8 + sizeof(S);
  return launch(s, PK_RESIDUAL, bpc * s->n, fsk::k_residual64<S>, grid2(s->nx, s->ny), BLK, (const float*)s->u[c],
                (const float*)s->v[c], (const double*)(zero_p ? nullptr : s->p64[pin]), e, s->p64[1 - pin],
                as<S>(s->lv[0].rhs), s->nx, s->ny, idx, idy, idx * idx, idy * idy, s->rpart + (size_t)it * s->rp_cols,
                s->gate_r);
}
int enqueue_residual(fs_solver_t* s, int it, int pin, bool zero_p, const float* e) {
  return s->bf16 ? enqueue_residual_s<fsk::bf16>(s, it, pin, zero_p, e) : enqueue_residual_s<float>(s, it, pin, zero_p, e);
}

// Residual `it`, then (unless residual_only) the V-cycle on its correction. split: the
// residual pass runs residual-only and the level-0 pre-smoothing is a k_mgt_down launch
// on the stored residual under the V-cycle's gate (the same bits), so that a pass that
// meets the gate, as a gated body's predicted last one usually does, skips its smoothing.
int enqueue_iteration(fs_solver_t* s, int it, int pin, bool zero_p, bool residual_only, bool split = false) {
  int rc;
  if (!s->fused) {
    if ((rc = enqueue_residual(s, it, pin, zero_p, it > 0 ? s->lv[0].x : nullptr))) return rc;
    return residual_only ? FS_OK : mg_vcycle(s);
  }
  int cs = s->coarse_start;
  if (cs == 0) {
    if ((rc = enqueue_residual(s, it, pin, zero_p, it > 0 ? s->lv[0].xb : nullptr))) return rc;
    return residual_only ? FS_OK : launch_coarse_ml(s, 0);
  }
  int pre = s->cfg.mg_pre, post = s->cfg.mg_post;
  if ((rc = DISPATCH_SW(pre, launch_down0, s, it, pin, zero_p, residual_only || split))) return rc;
  if (residual_only) return FS_OK;
  if (split && (rc = DISPATCH_SW(pre, launch_down, s, 0))) return rc;
  // the last two grid-wide levels above the coarse solve go down in one launch
  // (level l + 1 tiled by 8 x 8 exactly)
  const int pair = cs - 2 >= 1 && s->lv[cs - 1].n % 8 == 0 ? cs - 2 : cs;
  for (int l = 1; l < pair; l++)
    if ((rc = DISPATCH_SW(pre, launch_down, s, l))) return rc;
  if (pair < cs && (rc = DISPATCH_SW(pre, launch_down2, s, pair))) return rc;
  // the exact coarse solve runs in the first up launch when that launch is a k_mgt_up2
  const int up2_first = pair < cs && s->lv[pair].n % 16 == 0 ? pair : (cs == 2 && s->lv[0].n % 32 == 0 ? 0 : -1);
  if (!(up2_first >= 0 && fold_coarse(s, up2_first)) && (rc = launch_coarse_ml(s, cs))) return rc;
  // and come back up in one launch (level l tiled by 16 x 16 exactly); levels 1 and 0,
  // when neither is in that pair, also go up in one launch (32 x 32 tiles of level 0)
  int l = cs - 1;
  if (pair < cs && s->lv[pair].n % 16 == 0) {
    if ((rc = DISPATCH_SW(post, launch_up2, s, pair))) return rc;
    l = pair - 1;
  }
  for (; l >= 0; l--) {
    if (l == 1 && s->lv[0].n % 32 == 0) {
      if ((rc = DISPATCH_SW(post, launch_up2, s, 0))) return rc;
      break;
    }
    if ((rc = DISPATCH_SW(post, launch_up, s, l))) return rc;
  }
  return FS_OK;
}

int launch_gradient64(fs_solver_t* s, const fsk::FinalP* fp);

// The residual rows 0 .. rows-1 to the host, and their maxima; the gradient of the
// pressure `fp` selects is enqueued behind the copy and runs while the host waits for
// the copy alone.
int read_residuals_overlapped(fs_solver_t* s, int rows, std::vector<double>& rmax, const fsk::FinalP* fp) {
  HIPCHK(hipMemcpyAsync(s->rpart_host, s->rpart, sizeof(double) * (size_t)rows * s->rp_cols,
                        hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipEventRecord(s->rb_event, s->stream));
  int rc = launch_gradient64(s, fp);
  if (rc) return rc;
  HIPCHK(hipEventSynchronize(s->rb_event));
  after_sync(s);
  rmax.assign(rows, 0.0);
  for (int k = 0; k < rows; k++) {
    const double* row = s->rpart_host + (size_t)k * s->rp_cols;
    double m = 0;
    for (int b = 0; b < s->rp_cols; b++) m = std::max(m, row[b]);
    rmax[k] = m;
  }
  return FS_OK;
}

// ----------------------------------------------------------- learned guess

// Layer table of the network: input channels, output channels, kernel size, level
// (0 = n, 1 = n/2, ...), input mode and the channels that come from the coarser
// tensor when concatenating.
struct GuessLayer {
  int cin, cout, k, level, mode, c2;
};
static const GuessLayer kGuessLayers[15] = {
    {2, 16, 3, 0, fsk::CONV_DIRECT, 0},    // enc0a (VALU)
    {16, 16, 3, 0, fsk::CONV_DIRECT, 0},   // enc0b
    {16, 32, 3, 1, fsk::CONV_POOL, 0},     // enc1a
    {32, 32, 3, 1, fsk::CONV_DIRECT, 0},   // enc1b
    {32, 64, 3, 2, fsk::CONV_POOL, 0},     // enc2a
    {64, 64, 3, 2, fsk::CONV_DIRECT, 0},   // enc2b
    {64, 64, 3, 3, fsk::CONV_POOL, 0},     // bot_a
    {64, 64, 3, 3, fsk::CONV_DIRECT, 0},   // bot_b
    {128, 64, 3, 2, fsk::CONV_UPCAT, 64},  // dec2a = [up(bot_b), enc2b]
    {64, 64, 3, 2, fsk::CONV_DIRECT, 0},   // dec2b
    {96, 32, 3, 1, fsk::CONV_UPCAT, 64},   // dec1a = [up(dec2b), enc1b]
    {32, 32, 3, 1, fsk::CONV_DIRECT, 0},   // dec1b
    {48, 16, 3, 0, fsk::CONV_UPCAT, 32},   // dec0a = [up(dec1b), enc0b]
    {16, 16, 3, 0, fsk::CONV_DIRECT, 0},   // dec0b
    {16, 1, 1, 0, fsk::CONV_DIRECT, 0},    // out
};
// input tensors (x1, x2) of each MFMA layer by activation index (-1: none)
static const int kGuessIn[15][2] = {{-1, -1}, {0, -1}, {1, -1}, {2, -1}, {3, -1}, {4, -1}, {5, -1}, {6, -1},
                                    {5, 7},   {8, -1}, {3, 9},  {10, -1}, {1, 11}, {12, -1}, {13, -1}};

// Tile rows and output-channel blocks per workgroup, by level: large tiles where
// the level is big, output channels split across workgroups where it is small, so
// every level launches at least one workgroup per CU and the LDS holds the weights.
template <int CIN, int COUT, int MODE, int C2, int TH, int NBW, int WC = 1, int EPI = fsk::EPI_RELU>
int launch_conv(fs_solver_t* s, int layer, fsk::ConvEpiArgs e = fsk::ConvEpiArgs{}) {
  auto& G = s->guess_net;
  const int m = G.m >> kGuessLayers[layer].level;
  const uint16_t* x1 = G.act[kGuessIn[layer][0]];
  const uint16_t* x2 = kGuessIn[layer][1] >= 0 ? G.act[kGuessIn[layer][1]] : nullptr;
  // persistent workgroups: each stages its weight slice once and loops over tiles;
  // as many per CU as the LDS holds (at most 8), over all channel blocks
  const int ntiles = (m / fsk::GT_W) * (m / TH), nyb = COUT / (16 * NBW);
  const int lds = fsk::conv_lds_bytes<CIN, TH, NBW>();
  const int per_cu = std::max(1, std::min(8, (160 * 1024) / lds));
  dim3 g((unsigned)std::max(1, std::min(ntiles, 256 * per_cu / nyb)), (unsigned)nyb);
  double in_ch = MODE == fsk::CONV_POOL ? 4.0 * CIN : MODE == fsk::CONV_UPCAT ? C2 / 4.0 + (CIN - C2) : CIN;
  // activations in (pooled: 4 fine pixels per pixel; upsampled: 1 coarse pixel per 4) and
  // out, BF16; with the output layer fused, the FP64 pressure read and written instead
  const double out_bytes = EPI == fsk::EPI_OUT ? 16.0 : 2.0 * COUT;
  return launch(s, PK_GUESS, (2.0 * in_ch + out_bytes) * m * m,
                fsk::k_conv3x3<CIN, COUT, MODE, C2, TH, NBW, EPI, 0, WC>, g, dim3(fsk::GT_NT * WC), x1, x2,
                (const uint16_t*)G.w[layer], (const float*)G.b[layer], G.act[layer], m, 1, e);
}

// Evaluate the network on r (level-0 residual of p_prev, FP32) and p_prev = p64[pi],
// leaving p_prev + s_r h^2 out in p64[pi]. Inputs come from the last residual pass
// (rpart row 0 holds its per-tile maxima).
int run_guess(fs_solver_t* s, int pi) {
  auto& G = s->guess_net;
  const int n = s->nx;
  // FS_GUESS_LEARNED_COARSE / _QUARTER: the network on the 2^k x 2^k averages of r and
  // p_prev at n / 2^k (k = 1, 2), its increment interpolated bilinearly and added to
  // p_prev at n
  const int k = s->cfg.guess == FS_GUESS_LEARNED_COARSE ? 1 : s->cfg.guess == FS_GUESS_LEARNED_QUARTER ? 2 : 0;
  const bool coarse = k > 0;
  const int m = n >> k;
  G.m = m;
  int rc;
  if (coarse) {
    const int nb = m * m / 256;
    const double bytes = (s->bf16 ? 2.0 : 4.0) * n * n + 8.0 * n * n + 20.0 * m * m;
    if (s->bf16)
      rc = launch(s, PK_GUESS, bytes, fsk::k_guess_restrict<fsk::bf16>, dim3(nb), dim3(256),
                  (const fsk::bf16*)as<fsk::bf16>(s->lv[0].rhs), (const double*)s->p64[pi], n, k, G.rc, G.pc, G.dc,
                  G.cpart);
    else
      rc = launch(s, PK_GUESS, bytes, fsk::k_guess_restrict<float>, dim3(nb), dim3(256), (const float*)s->lv[0].rhs,
                  (const double*)s->p64[pi], n, k, G.rc, G.pc, G.dc, G.cpart);
    if (rc) return rc;
    fsk::k_guess_scales<<<1, 256, 0, s->stream>>>(G.cpart, nb, G.cpart + nb, nb, G.scales);
    LAUNCHCHK();
  } else {
    fsk::k_absmax64<<<G.nblk, 256, 0, s->stream>>>(s->p64[pi], s->n, G.ppart);
    LAUNCHCHK();
    fsk::k_guess_scales<<<1, 256, 0, s->stream>>>(s->rpart, s->rp_cols, G.ppart, G.nblk, G.scales);
    LAUNCHCHK();
  }
  // enc0a and enc0b in one kernel (enc0a's output stays in LDS); the residual r is read
  // at its storage precision (at half resolution: the FP32 averages)
  {
    constexpr int TH0 = 8;
    const int ntiles = (m / fsk::GT_W) * (m / TH0);
    const dim3 g((unsigned)std::min(ntiles, 256 * 8));
    const bool rb = s->bf16 && !coarse;
    const double bytes = (rb ? 2.0 : 4.0) * m * m + 8.0 * m * m + 32.0 * m * m;
    const double* pin = coarse ? G.pc : s->p64[pi];
    if (rb)
      rc = launch(s, PK_GUESS, bytes, fsk::k_guess_enc0<fsk::bf16, TH0>, g, dim3(fsk::GT_NT),
                  (const fsk::bf16*)as<fsk::bf16>(s->lv[0].rhs), pin, (const double*)G.scales, (const float*)G.w0,
                  (const float*)G.b[0], (const uint16_t*)G.w[1], (const float*)G.b[1], G.act[1], m);
    else
      rc = launch(s, PK_GUESS, bytes, fsk::k_guess_enc0<float, TH0>, g, dim3(fsk::GT_NT),
                  coarse ? (const float*)G.rc : (const float*)s->lv[0].rhs, pin, (const double*)G.scales,
                  (const float*)G.w0, (const float*)G.b[0], (const uint16_t*)G.w[1], (const float*)G.b[1], G.act[1],
                  m);
    if (rc) return rc;
  }
  // Tile rows, channel blocks and wavefront columns per layer: deep tiles where the
  // weights are small, channels split over workgroups or wavefronts where input +
  // weights would crowd the LDS. Per layer, by its output size (a network at n / 2 or n / 4 makes every level
  // smaller; timed at network inputs of 1024, 512 and 256, all shapes giving the same
  // bits): short tiles and fewer channel blocks per workgroup on the small levels,
  // where more workgroups in flight is what pays.
  const int m1 = m >> 1, m2 = m >> 2, m3 = m >> 3;
  using namespace fsk;
  if ((rc = launch_conv<16, 32, CONV_POOL, 0, 8, 2, 2>(s, 2))) return rc;
  if ((rc = m1 <= 128   ? launch_conv<32, 32, CONV_DIRECT, 0, 4, 2>(s, 3)
            : m1 <= 256 ? launch_conv<32, 32, CONV_DIRECT, 0, 8, 2>(s, 3)
                        : launch_conv<32, 32, CONV_DIRECT, 0, 16, 2>(s, 3)))
    return rc;
  if ((rc = m2 <= 64    ? launch_conv<32, 64, CONV_POOL, 0, 4, 2, 2>(s, 4)
            : m2 <= 128 ? launch_conv<32, 64, CONV_POOL, 0, 8, 2, 2>(s, 4)
                        : launch_conv<32, 64, CONV_POOL, 0, 8, 4, 2>(s, 4)))
    return rc;
  if ((rc = m2 <= 64    ? launch_conv<64, 64, CONV_DIRECT, 0, 4, 1>(s, 5)
            : m2 <= 128 ? launch_conv<64, 64, CONV_DIRECT, 0, 8, 2, 2>(s, 5)
                        : launch_conv<64, 64, CONV_DIRECT, 0, 16, 4, 2>(s, 5)))
    return rc;
  if ((rc = m3 <= 32    ? launch_conv<64, 64, CONV_POOL, 0, 4, 1>(s, 6)
            : m3 <= 64  ? launch_conv<64, 64, CONV_POOL, 0, 4, 2, 2>(s, 6)
                        : launch_conv<64, 64, CONV_POOL, 0, 8, 2, 2>(s, 6)))
    return rc;
  if ((rc = m3 <= 64 ? launch_conv<64, 64, CONV_DIRECT, 0, 4, 1>(s, 7)
                     : launch_conv<64, 64, CONV_DIRECT, 0, 16, 1>(s, 7)))
    return rc;
  if ((rc = m2 <= 64    ? launch_conv<128, 64, CONV_UPCAT, 64, 8, 1>(s, 8)
            : m2 <= 128 ? launch_conv<128, 64, CONV_UPCAT, 64, 8, 2, 2>(s, 8)
                        : launch_conv<128, 64, CONV_UPCAT, 64, 16, 2, 2>(s, 8)))
    return rc;
  if ((rc = m2 <= 64    ? launch_conv<64, 64, CONV_DIRECT, 0, 4, 1>(s, 9)
            : m2 <= 128 ? launch_conv<64, 64, CONV_DIRECT, 0, 8, 2, 2>(s, 9)
                        : launch_conv<64, 64, CONV_DIRECT, 0, 16, 4, 2>(s, 9)))
    return rc;
  if ((rc = m1 <= 128 ? launch_conv<96, 32, CONV_UPCAT, 64, 8, 1>(s, 10)
                      : launch_conv<96, 32, CONV_UPCAT, 64, 16, 2, 2>(s, 10)))
    return rc;
  if ((rc = m1 <= 128   ? launch_conv<32, 32, CONV_DIRECT, 0, 4, 1>(s, 11)
            : m1 <= 256 ? launch_conv<32, 32, CONV_DIRECT, 0, 8, 2>(s, 11)
                        : launch_conv<32, 32, CONV_DIRECT, 0, 16, 2>(s, 11)))
    return rc;
  if ((rc = launch_conv<48, 16, CONV_UPCAT, 32, 16, 1>(s, 12))) return rc;
  // the last 3x3 layer with the 1x1 output layer and the pressure update in its epilogue
  fsk::ConvEpiArgs out{};
  out.wout = G.wout;
  out.bout = G.b[14];
  out.scales = G.scales;
  out.p = coarse ? G.dc : s->p64[pi];
  out.h2 = (double)(1 << (2 * k)) * s->dx * s->dx;
  if ((rc = m == 512 ? launch_conv<16, 16, CONV_DIRECT, 0, 16, 1, 1, EPI_OUT>(s, 13, out)
                     : launch_conv<16, 16, CONV_DIRECT, 0, 8, 1, 1, EPI_OUT>(s, 13, out)))
    return rc;
  if (coarse)
    rc = launch(s, PK_GUESS, 8.0 * m * m + 16.0 * n * n, fsk::k_guess_prolong_add, dim3(n * n / 512), dim3(256),
                (const double*)G.dc, s->p64[pi], n, k);
  return rc;
}

void release_guess(fs_solver_t* s) {
  auto& G = s->guess_net;
  dfree(G.w0);
  dfree(G.wout);
  for (int l = 0; l < 15; l++) {
    dfree(G.b[l]);
    dfree(G.w[l]);
  }
  for (int a = 0; a < 14; a++) dfree(G.act[a]);
  dfree(G.scales);
  dfree(G.ppart);
  dfree(G.rc);
  dfree(G.pc);
  dfree(G.dc);
  dfree(G.cpart);
  G = fs_solver_t::GuessNet{};
}

bool learned_guess(const fs_solver_t* s);

// Whether the projection starts from a zero pressure (no warm start configured or none yet).
bool zero_start(const fs_solver_t* s) {
  const bool warm = s->cfg.guess == FS_GUESS_PREVIOUS || s->cfg.guess == FS_GUESS_EXTRAPOLATED || learned_guess(s);
  return !(warm && s->have_p);
}

// The initial pressure before the refinement iterations: the extrapolation in time
// (FS_GUESS_EXTRAPOLATED), then the network's increment (the learned guesses) on top of
// the previous pressure. Moves a0 to the buffer the guess is in.
int enqueue_guess(fs_solver_t* s, int& a0, bool& zero_p) {
  int rc;
  if (s->cfg.guess == FS_GUESS_EXTRAPOLATED && !zero_p) {
    const int out = 1 - (a0 & 1);
    fsk::k_p_extrapolate<<<(unsigned)((s->n / 2 + 255) / 256), 256, 0, s->stream>>>(
        s->p64[a0 & 1], s->p_hist, s->p64[out], s->n, s->have_hist ? 1 : 0);
    LAUNCHCHK();
    s->have_hist = true;
    a0 = out;
  }
  if (learned_guess(s)) {
    if (!s->guess_net.loaded) return fail(FS_E_INVALID_ARG, "the learned guess has no weights loaded");
    // the residual the previous pressure leaves, the network's increment on top of it,
    // then refinement from that guess
    if ((rc = enqueue_iteration(s, 0, a0 & 1, zero_p, true))) return rc;
    if ((rc = run_guess(s, 1 - (a0 & 1)))) return rc;
    a0 = 1 - (a0 & 1);
    zero_p = false;
  }
  return FS_OK;
}

// Multigrid projection by iterative refinement; leaves the pressure in p64[s->pidx].
int solve_multigrid(fs_solver_t* s, int* cycles_out) {
  const double tol = s->cfg.tolerance;
  const int maxc = std::max(0, s->cfg.pressure_iterations);
  bool zero_p = zero_start(s);
  int a0 = s->pidx;
  int rc;
  if ((rc = enqueue_guess(s, a0, zero_p))) return rc;
  if (!(tol > 0)) {
    // k V-cycles: residual(0), V(0), ..., residual(k-1), V(k-1), residual(k); no host round trip
    for (int it = 0; it <= maxc; it++)
      if ((rc = enqueue_iteration(s, it, (a0 + it) & 1, it == 0 && zero_p, it == maxc))) return rc;
    HIPCHK(hipMemcpyAsync(s->rpart_host + (size_t)maxc * s->rp_cols, s->rpart + (size_t)maxc * s->rp_cols,
                          sizeof(double) * s->rp_cols, hipMemcpyDeviceToHost, s->stream));
    s->resid_row_pending = maxc;
    s->pidx = (a0 + maxc + 1) & 1;
    *cycles_out = maxc;
    return FS_OK;
  }
  // Predicted V-cycles, then one read-back; add V-cycles one at a time if needed.
  // On the fused path every launch is gated on the device as in a gated step, so the
  // projection stops at the first iteration that meets the gate even when the
  // prediction was larger; rows after that iteration are stale and never consulted.
  // The per-operator path has no gated launches and searches from zero instead.
  // While the host waits for the residuals, the device already subtracts the gradient
  // of the pressure they belong to (into the other velocity buffer); the projection
  // keeps it if the gate is met and enqueues it again otherwise.
  int c = s->fused ? std::max(0, std::min(s->last_pred, maxc)) : 0;
  if (s->fused && (rc = fill_f32(s, (float*)s->gate_flag, (size_t)s->gate_len, 0.f))) return rc;
  auto iteration = [&](int it, bool residual_only) {
    if (s->fused) {
      s->gate_r = fsk::Gate{s->gate_flag, tol, it > 0 ? it - 1 : -1, it};
      s->gate_v = fsk::Gate{s->gate_flag, tol, it, -1};
    }
    int r = enqueue_iteration(s, it, (a0 + it) & 1, it == 0 && zero_p, residual_only);
    s->gate_r = s->gate_v = fsk::gate_open();
    return r;
  };
  for (int it = 0; it < c; it++)
    if ((rc = iteration(it, false))) return rc;
  if ((rc = iteration(c, true))) return rc;
  std::vector<double> rmax;
  for (;;) {
    s->pidx = (a0 + c + 1) & 1;
    fsk::FinalP f{};
    f.flag = s->gate_flag;
    f.k = s->fused ? c : -1;
    f.a0 = a0;
    f.keep = s->pidx;
    f.pbuf[0] = s->fused ? s->p64[0] : s->p64[s->pidx];
    f.pbuf[1] = s->fused ? s->p64[1] : s->p64[s->pidx];
    if ((rc = read_residuals_overlapped(s, c + 1, rmax, &f))) return rc;
    int need = -1;
    for (int k = 0; k <= c && need < 0; k++) {
      if (std::isnan(rmax[k]) || std::isinf(rmax[k])) {
        *cycles_out = k;
        s->last_residual = rmax[k];
        s->have_maxabs = false;  // the gradient pass that ran rewrote the |u|, |v| maxima
        return fail(FS_E_NONFINITE, "non-finite residual in the projection (V-cycle %d)", k);
      }
      if (rmax[k] <= tol) need = k;
    }
    if (need >= 0 || c >= maxc) {
      const int used = need >= 0 ? need : c;
      s->grad_pending = true;
      s->last_pred = used;
      s->pred_known = true;
      s->last_residual = rmax[used];
      *cycles_out = used;
      if (need < 0)
        return fail(FS_E_NOT_CONVERGED, "multigrid did not reach tolerance %.3g in %d V-cycles (residual %.3g)",
                    tol, maxc, rmax[c]);
      return FS_OK;
    }
    // residual c was enqueued residual-only: rerun it with its V-cycle, then probe c+1
    if ((rc = iteration(c, false))) return rc;
    c++;
    if ((rc = iteration(c, true))) return rc;
  }
}

// u - grad p into the other velocity buffer, with the |u|, |v| block maxima; the
// solver's state is not changed (see commit_gradient64). Without a FinalP the
// pressure is p64[pidx].
int launch_gradient64(fs_solver_t* s, const fsk::FinalP* fp) {
  int c = s->cur;
  double idx = 1.0 / s->dx, idy = 1.0 / s->dy;
  fsk::FinalP f{};
  if (fp) {
    f = *fp;
  } else {
    f.k = -1;
    f.pbuf[0] = f.pbuf[1] = s->p64[s->pidx];
  }
  return launch(s, PK_GRADIENT, 24.0 * s->n, fsk::k_gradient64, grid2(s->nx, s->ny), dim3(64), (const float*)s->u[c],
                (const float*)s->v[c], f, s->u[1 - c], s->v[1 - c], s->nx, s->ny, idx, idy, s->mpart);
}

void commit_gradient64(fs_solver_t* s) {
  s->have_maxabs = true;
  s->cur = 1 - s->cur;
}

int enqueue_gradient64(fs_solver_t* s) {
  int rc = launch_gradient64(s, nullptr);
  if (rc) return rc;
  commit_gradient64(s);
  return FS_OK;
}

int enqueue_projection(fs_solver_t* s, int* cycles) {
  *cycles = 0;
  if (s->cfg.pressure_solver == FS_P_MULTIGRID) {
    s->grad_pending = false;
    int rc = solve_multigrid(s, cycles);
    if (rc && rc != FS_E_NOT_CONVERGED) return rc;
    s->have_p = true;
    if (s->grad_pending) {  // enqueued while the residuals were read
      s->grad_pending = false;
      commit_gradient64(s);
      return rc;
    }
    int rc2 = enqueue_gradient64(s);
    return rc2 ? rc2 : rc;
  }
  int rc = enqueue_divergence(s);
  if (rc) return rc;
  int pi = 0;
  if ((rc = solve_flat(s, s->cfg.pressure_iterations, &pi))) return rc;
  s->p32idx = pi;
  {
    int c = s->cur;
    fsk::k_gradient32<<<grid2(s->nx, s->ny), BLK, 0, s->stream>>>(s->u[c], s->v[c], s->p32[pi], s->u[1 - c],
                                                                    s->v[1 - c], s->nx, s->ny, (float)(1.0 / s->dx),
                                                                    (float)(1.0 / s->dy));
    LAUNCHCHK();
    s->cur = 1 - c;
  }
  s->have_maxabs = false;
  return FS_OK;
}

// The step after its coefficients were enqueued.
int do_step(fs_solver_t* s) {
  int rc;
  if (s->dt_pending >= DT_RING && (rc = sync_stream(s))) return rc;
  HIPCHK(hipMemcpyAsync(s->dt_ring + s->dt_pending, &s->coef->dt, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  s->dt_pending++;
  if (s->cfg.advection == FS_ADV_MACCORMACK && s->cfg.adv_fused) {
    if ((rc = enqueue_advect_diffuse_fused(s))) return rc;
  } else {
    if ((rc = enqueue_advection(s))) return rc;
    if ((rc = enqueue_diffusion(s))) return rc;
  }
  if ((rc = enqueue_forcing(s))) return rc;
  if ((rc = enqueue_splats(s))) return rc;
  if ((rc = enqueue_confinement(s))) return rc;
  if ((rc = enqueue_solid(s))) return rc;
  int cycles = 0;
  int prc = enqueue_projection(s, &cycles);
  if (prc && prc != FS_E_NOT_CONVERGED) return prc;
  if ((rc = enqueue_dye(s))) return rc;
  s->last_vcycles = cycles;
  s->total_vcycles += cycles;
  s->steps++;
  return prc;
}


// ------------------------------------------------------- device-gated step

// Read the records of gated steps whose event has completed, oldest first; with
// `wait`, wait for all of them. The prediction for the next steps is the V-cycle
// count of the latest record (one more than it ran when it missed the gate).
int consume_records(fs_solver_t* s, bool wait) {
  size_t done = 0;
  int rc = FS_OK;
  for (; done < s->inflight.size(); done++) {
    const auto& f = s->inflight[done];
    if (wait) {
      hipError_t e = hipEventSynchronize(f.ev);
      if (e != hipSuccess) {
        rc = fail(FS_E_DEVICE, "hipEventSynchronize failed: %s", hipGetErrorString(e));
        break;
      }
    } else {
      hipError_t e = hipEventQuery(f.ev);
      if (e == hipErrorNotReady) break;
      if (e != hipSuccess) {
        rc = fail(FS_E_DEVICE, "hipEventQuery failed: %s", hipGetErrorString(e));
        break;
      }
    }
    const fsk::StepRecord r = s->ring[f.step % REC_RING];
    s->ev_pool.push_back(f.ev);
    if (r.step != f.step) {
      done++;
      rc = fail(FS_E_DEVICE, "the record of step %lld is missing", (long long)f.step);
      break;
    }
    s->t += r.dt;
    s->last_dt = r.dt;
    s->last_vcycles = r.vcycles;
    s->total_vcycles += r.vcycles;
    s->last_residual = r.residual;
    s->last_pred = r.missed ? r.vcycles + 1 : r.vcycles;
    if (r.missed) s->gate_misses++;
    if (!std::isfinite(r.residual) && rc == FS_OK)
      rc = fail(FS_E_NONFINITE, "non-finite residual in the projection of step %lld", (long long)f.step);
  }
  s->inflight.erase(s->inflight.begin(), s->inflight.begin() + done);
  return rc;
}

void destroy_graphs(fs_solver_t* s) {
  if (s->graphs.empty()) return;
  (void)hipStreamSynchronize(s->stream);
  for (auto& g : s->graphs) (void)hipGraphExecDestroy(g.exec);
  s->graphs.clear();
}

bool learned_guess(const fs_solver_t* s) {
  return s->cfg.guess == FS_GUESS_LEARNED || s->cfg.guess == FS_GUESS_LEARNED_COARSE ||
         s->cfg.guess == FS_GUESS_LEARNED_QUARTER;
}

// A tolerance-driven multigrid projection on the fused V-cycle, with a prediction
// from the current flow and no launch timing (a profiled pass times every launch it
// enqueues, so it takes the host-checked path).
bool gated_eligible(const fs_solver_t* s) {
  return s->gated && s->pred_known && s->cfg.pressure_solver == FS_P_MULTIGRID && s->cfg.tolerance > 0 &&
         s->fused && s->gate_flag && s->prof.mask == 0;
}

// What the launches of a gated step body depend on besides the configuration:
// iterations k, the pressure buffer the step starts from, a zero initial pressure,
// the velocity buffer, and the buffer the pressure is left in.
struct BodyKey {
  int k, a0, zero_p, cur, keep, hist, dye;
};

// The step after its head: advection, diffusion, forcing, disk; the learned guess;
// refinement iterations 0 .. k-1, each launch gated on the flags of its residual
// pass; the residual pass k; the gradient of the pressure of the first iteration
// that met the gate, left in p64[keep]. Launches only, so it can be captured.
int enqueue_gated_body(fs_solver_t* s, const BodyKey& key) {
  int rc;
  if (s->cfg.advection == FS_ADV_MACCORMACK && s->cfg.adv_fused) {
    if ((rc = enqueue_advect_diffuse_fused(s))) return rc;
  } else {
    if ((rc = enqueue_advection(s))) return rc;
    if ((rc = enqueue_diffusion(s))) return rc;
  }
  if ((rc = enqueue_forcing(s))) return rc;
  if ((rc = enqueue_splats(s))) return rc;
  if ((rc = enqueue_confinement(s))) return rc;
  if ((rc = enqueue_solid(s))) return rc;
  int a0 = key.a0;
  bool zero_p = key.zero_p != 0;
  const bool hist0 = s->have_hist;
  if ((rc = enqueue_guess(s, a0, zero_p))) return rc;
  if (key.hist != (hist0 ? 1 : 0)) return fail(FS_E_DEVICE, "step body key out of date");
  const double tol = s->cfg.tolerance;
  // iteration k - 1's residual pass is the predicted last one, and its V-cycle the spare
  for (int it = 0; it <= key.k && !rc; it++) {
    s->gate_r = fsk::Gate{s->gate_flag, tol, it > 0 ? it - 1 : -1, it};
    s->gate_v = fsk::Gate{s->gate_flag, tol, it, -1};
    s->spare = it == key.k - 1;
    rc = enqueue_iteration(s, it, (a0 + it) & 1, it == 0 && zero_p, it == key.k, it == key.k - 1);
    s->spare = false;
  }
  s->gate_r = s->gate_v = fsk::gate_open();
  if (rc) return rc;
  fsk::FinalP f{};
  f.flag = s->gate_flag;
  f.rpart = s->rpart;
  f.cols = s->rp_cols;
  f.k = key.k;
  f.a0 = a0;
  f.keep = key.keep;
  f.pbuf[0] = s->p64[0];
  f.pbuf[1] = s->p64[1];
  f.ring = s->ring;
  f.ring_len = REC_RING;
  f.coef = s->coef;
  if ((rc = launch_gradient64(s, &f))) return rc;
  commit_gradient64(s);
  return enqueue_dye(s);
}

const fs_solver_t::StepGraph* find_graph(const fs_solver_t* s, const BodyKey& key) {
  for (const auto& g : s->graphs)
    if (g.k == key.k && g.a0 == key.a0 && g.zero_p == key.zero_p && g.cur == key.cur && g.keep == key.keep &&
        g.hist == key.hist && g.dye == key.dye)
      return &g;
  return nullptr;
}

// Capture the body of `key` into a new cached graph without launching it. Capturing
// enqueues nothing: the solver state the enqueue functions advance (velocity and dye
// parity, the flags) is put back afterwards.
int capture_body(fs_solver_t* s, const BodyKey& key) {
  const int cur0 = s->cur, dye0 = s->dye_cur;
  const bool maxabs0 = s->have_maxabs, hist0 = s->have_hist;
  HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_gated_body(s, key);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(s->stream, &graph);
  const int cur_after = s->cur, dye_after = s->dye_cur;
  s->cur = cur0, s->dye_cur = dye0, s->have_maxabs = maxabs0, s->have_hist = hist0;
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? rc : fail(FS_E_DEVICE, "capturing the step failed: %s", hipGetErrorString(e));
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(FS_E_DEVICE, "instantiating the step graph failed: %s", hipGetErrorString(e));
  s->graphs.push_back({key.k, key.a0, key.zero_p, key.cur, key.keep, key.hist, key.dye, cur_after, dye_after, exec});
  s->graph_captures++;
  return FS_OK;
}

// The body as a graph, captured on first use of its key (unless captured ahead) and
// replayed after.
int launch_gated_body_graph(fs_solver_t* s, const BodyKey& key) {
  const fs_solver_t::StepGraph* g = find_graph(s, key);
  if (!g) {
    if (s->graphs.size() >= 32) destroy_graphs(s);  // a drifting prediction: start afresh
    int rc = capture_body(s, key);
    if (rc) return rc;
    g = &s->graphs.back();
  }
  HIPCHK(hipGraphLaunch(g->exec, s->stream));
  s->graph_launches++;
  s->cur = g->cur_after;
  s->dye_cur = g->dye_after;
  s->have_maxabs = true;
  if (s->cfg.guess == FS_GUESS_EXTRAPOLATED && !key.zero_p) s->have_hist = true;
  return FS_OK;
}

// The key of the next gated step body if its V-cycle prediction were `pred`.
BodyKey gated_key(const fs_solver_t* s, int pred) {
  const int maxc = std::max(0, s->cfg.pressure_iterations);
  pred = std::max(0, std::min(pred, maxc));
  BodyKey key{};
  key.k = std::min(pred + GATE_SPARE, maxc);
  key.a0 = s->pidx;
  key.zero_p = zero_start(s) ? 1 : 0;
  key.cur = s->cur;
  key.hist = s->have_hist ? 1 : 0;
  key.dye = s->dye[0] ? s->dye_cur : -1;
  int a0e = s->pidx;
  if (s->cfg.guess == FS_GUESS_EXTRAPOLATED && !key.zero_p) a0e = 1 - (a0e & 1);
  if (learned_guess(s)) a0e = 1 - (a0e & 1);
  key.keep = (a0e + pred + 1) & 1;
  return key;
}

// After a step is enqueued: capture one body the next step may need and that is not
// cached yet (its current prediction, then one V-cycle more, then one fewer). A
// capture on demand leaves the device idle while the host records and instantiates
// it; ahead of time it overlaps the steps already queued.
int capture_ahead(fs_solver_t* s) {
  if (s->graphs.size() >= 24) return FS_OK;
  for (int d : {0, 1, -1}) {
    const int pred = s->last_pred + d;
    if (pred < 0 || pred > s->cfg.pressure_iterations) continue;
    const BodyKey key = gated_key(s, pred);
    if (!find_graph(s, key)) {
      (void)capture_body(s, key);  // speculative: a failure leaves the capture to the step that needs it
      return FS_OK;
    }
  }
  return FS_OK;
}

// One step whose projection is checked on the device: the predicted V-cycles plus a
// spare are enqueued, the host reads the step's record once its event has completed.
int step_gated(fs_solver_t* s, bool use_cfl, double value) {
  int rc = consume_records(s, false);
  while (rc == FS_OK && (int)s->inflight.size() >= MAX_AHEAD) {
    // bound how far the host runs ahead of the device (and how stale the prediction gets)
    HIPCHK(hipEventSynchronize(s->inflight.front().ev));
    rc = consume_records(s, false);
  }
  if (rc) return rc;
  if (learned_guess(s) && !s->guess_net.loaded)
    return fail(FS_E_INVALID_ARG, "the learned guess has no weights loaded");
  const BodyKey key = gated_key(s, s->last_pred);
  if ((rc = enqueue_head(s, use_cfl, value, key.k + 1))) return rc;
  const bool graph = s->use_graph && !s->solid_on && s->splats.empty();
  if (graph) rc = launch_gated_body_graph(s, key);
  else rc = enqueue_gated_body(s, key);
  if (rc) return rc;
  hipEvent_t ev = nullptr;
  if (!s->ev_pool.empty()) {
    ev = s->ev_pool.back();
    s->ev_pool.pop_back();
  } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
    return fail(FS_E_DEVICE, "hipEventCreate failed");
  }
  HIPCHK(hipEventRecord(ev, s->stream));
  s->inflight.push_back({ev, s->steps});
  s->pidx = key.keep;
  s->have_p = true;
  s->steps++;
  s->gated_steps++;
  return graph ? capture_ahead(s) : FS_OK;
}


// The coarsest level's red-black sweeps converge geometrically with factor
// rho = ((1 + cos(2 pi / n)) / 2)^2 (the smoothest non-constant mode). When
// rho^sweeps is below FP32 rounding the sweeps are an exact solve, and the
// pseudo-inverse of the periodic 5-point Laplacian gives the same answer (up to
// the null-space constant, which no residual sees) in one pass. Built in FP64:
// A+ = (A + 11^T/N)^-1 - 11^T/N.
int build_coarse_inverse(fs_solver_t* s) {
  const Level& C = s->lv.back();
  int n = C.n, N = n * n;
  if (n > fsk::REG_N || n < 2) return FS_OK;
  double rho = std::pow((1.0 + std::cos(2 * M_PI / n)) / 2.0, 2.0);
  if (!(std::pow(rho, (double)s->cfg.mg_coarse) < 1e-7)) return FS_OK;
  double h2 = C.dx * C.dx;
  std::vector<double> M((size_t)N * 2 * N, 0.0);  // [A + 11^T/N | I]
  auto at = [&](int r, int c) -> double& { return M[(size_t)r * 2 * N + c]; };
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      int r = j * n + i;
      int nb[4] = {j * n + (i + n - 1) % n, j * n + (i + 1) % n, ((j + n - 1) % n) * n + i, ((j + 1) % n) * n + i};
      at(r, r) += -4.0 / h2;
      for (int q = 0; q < 4; q++) at(r, nb[q]) += 1.0 / h2;
      for (int c = 0; c < N; c++) at(r, c) += 1.0 / N;
      at(r, N + r) = 1.0;
    }
  for (int col = 0; col < N; col++) {  // Gauss-Jordan with partial pivoting
    int piv = col;
    for (int r = col + 1; r < N; r++)
      if (std::fabs(at(r, col)) > std::fabs(at(piv, col))) piv = r;
    if (std::fabs(at(piv, col)) < 1e-300) return fail(FS_E_DEVICE, "singular coarse-level matrix");
    if (piv != col)
      for (int c = 0; c < 2 * N; c++) std::swap(at(piv, c), at(col, c));
    double d = at(col, col);
    for (int c = 0; c < 2 * N; c++) at(col, c) /= d;
    for (int r = 0; r < N; r++)
      if (r != col) {
        double f = at(r, col);
        if (f != 0.0)
          for (int c = 0; c < 2 * N; c++) at(r, c) -= f * at(col, c);
      }
  }
  // stored transposed: pinv[c*N + r] = A+(r, c), so lanes read consecutive words
  std::vector<float> pt((size_t)N * N);
  for (int r = 0; r < N; r++)
    for (int c = 0; c < N; c++) pt[(size_t)c * N + r] = (float)(at(r, N + c) - 1.0 / N);
  int rc = dalloc(&s->pinv, (size_t)N * N);
  if (rc) return rc;
  HIPCHK(hipMemcpy(s->pinv, pt.data(), sizeof(float) * N * N, hipMemcpyHostToDevice));
  return FS_OK;
}

// Table of the exact coarsest-level solve (k_mg_coarse_exact), built in FP64 for an
// n x n level of spacing h: an orthonormal eigenbasis Q of the 1D second difference and
// W[a][b] = 1 / (lam_a + lam_b) with W[0][0] = 0, zero-padded to 64 x 64. Periodic: the
// real Fourier basis (column 0 the constant, columns 2f-1 and 2f the cosine and sine of
// frequency f, column n-1 the alternating mode when n is even), lam of frequency f =
// -(4 / h^2) sin^2(pi f / n). Walls (cell-centred, zero flux): the DCT-II basis
// cos(pi k (j + 1/2) / n), lam_k = -(4 / h^2) sin^2(pi k / (2 n)). Q goes in as BF16
// heads and tails in the order the lanes read them, W as FP32 at each lane's outputs
// (fsk::EX_TAB 16-byte units).
uint16_t bf16_bits(double x) {
  float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even (finite input)
  return (uint16_t)(u >> 16);
}
double bf16_value(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

std::vector<uint8_t> exact_coarse_table(int n, double h, bool walls) {
  const int E = fsk::EXN;
  std::vector<double> Q((size_t)E * E, 0.0), lam(E, 0.0);
  for (int c = 0; c < n; c++) {
    if (walls) {
      const double sh = std::sin(M_PI * c / (2.0 * n));
      lam[c] = -4.0 * sh * sh / (h * h);
      const double a = c == 0 ? std::sqrt(1.0 / n) : std::sqrt(2.0 / n);
      for (int j = 0; j < n; j++) Q[(size_t)j * E + c] = a * std::cos(M_PI * c * (j + 0.5) / n);
      continue;
    }
    const int f = (c + 1) / 2;  // 0; 1, 1; 2, 2; ...; n/2 for the last column when n is even
    const bool alt = n % 2 == 0 && c == n - 1, sine = c > 0 && c % 2 == 0 && !alt;
    const double sh = std::sin(M_PI * f / n);
    lam[c] = -4.0 * sh * sh / (h * h);
    for (int j = 0; j < n; j++) {
      double v;
      if (c == 0) v = 1.0 / std::sqrt((double)n);
      else if (alt) v = ((j & 1) ? -1.0 : 1.0) / std::sqrt((double)n);
      else if (sine) v = std::sqrt(2.0 / n) * std::sin(2 * M_PI * f * j / n);
      else v = std::sqrt(2.0 / n) * std::cos(2 * M_PI * f * j / n);
      Q[(size_t)j * E + c] = v;
    }
  }
  std::vector<uint8_t> out((size_t)fsk::EX_TAB * 16);
  uint16_t* qt = (uint16_t*)out.data();
  for (int T = 0; T < 4; T++)
    for (int lane = 0; lane < 64; lane++)
      for (int kk = 0; kk < 2; kk++)
        for (int t = 0; t < 8; t++) {
          const int li = lane & 15, k = 32 * kk + 8 * (lane >> 4) + t;
          const size_t o = ((size_t)(T * 64 + lane) * 2 + kk) * 8 + t;
          const double a = Q[(size_t)k * E + 16 * T + li];  // Q[k][16 T + i]
          const double b = Q[(size_t)(16 * T + li) * E + k];  // Q[16 T + i][k]
          const uint16_t ah = bf16_bits(a), bh = bf16_bits(b);
          qt[o] = ah;
          qt[512 * 8 + o] = bf16_bits(a - bf16_value(ah));
          qt[2 * 512 * 8 + o] = bh;
          qt[3 * 512 * 8 + o] = bf16_bits(b - bf16_value(bh));
        }
  float* wt = (float*)(out.data() + 4 * 512 * 8 * 2);
  for (int w = 0; w < 16; w++)
    for (int lane = 0; lane < 64; lane++)
      for (int i = 0; i < 4; i++) {
        const int a = 16 * (w >> 2) + 4 * (lane >> 4) + i, b = 16 * (w & 3) + (lane & 15);
        wt[((size_t)w * 64 + lane) * 4 + i] =
            (a >= n || b >= n || (a == 0 && b == 0)) ? 0.f : (float)(1.0 / (lam[a] + lam[b]));
      }
  static_assert(fsk::EX_TAB * 16 == 4 * 512 * 8 * 2 + 16 * 64 * 4 * 4, "table layout");
  return out;
}

int build_coarse_exact(fs_solver_t* s) {
  const Level& C = s->lv.back();
  if (C.n > fsk::EXN) return fail(FS_E_INVALID_ARG, "exact coarse solve of a %d-cell level (at most %d)", C.n, fsk::EXN);
  const std::vector<uint8_t> t = exact_coarse_table(C.n, C.dx, false);
  int rc = dalloc(&s->ex_q, (size_t)fsk::EX_TAB * 4);
  if (rc) return rc;
  HIPCHK(hipMemcpy(s->ex_q, t.data(), t.size(), hipMemcpyHostToDevice));
  return FS_OK;
}

int build_levels(fs_solver_t* s) {
  if (s->cfg.pressure_solver != FS_P_MULTIGRID) return FS_OK;
  if (s->nx != s->ny) return fail(FS_E_INVALID_ARG, "multigrid here assumes a square grid, got %dx%d", s->nx, s->ny);
  int minS = s->cfg.mg_min_size;
  if (minS < 2) return fail(FS_E_INVALID_ARG, "mg_min_size must be >= 2, got %d", minS);
  int n = s->nx;
  if (n < minS) return fail(FS_E_INVALID_ARG, "grid %d is smaller than mg_min_size %d", n, minS);
  double L = s->cfg.lx;
  // src/gpu/multigrid.js:182-197: halve while n >= minSize. Every level stays even:
  // on a periodic grid of odd size the red-black colouring meets itself at the seam.
  while (n >= minS) {
    Level lvl;
    lvl.n = n;
    lvl.dx = L / n;
    s->lv.push_back(lvl);
    if (s->cfg.mg_coarse_exact && n <= fsk::COARSE_N) break;  // solved exactly: no levels below
    if (n / 2 < minS || (n & 1) || ((n / 2) & 1)) break;
    n /= 2;
  }
  if (s->cfg.mg_coarse_exact && s->lv.back().n > fsk::COARSE_N)
    return fail(FS_E_INVALID_ARG, "the exact coarse solve needs a level of at most %d cells; the hierarchy of %d "
                                  "with mg_min_size %d stops at %d", fsk::COARSE_N, s->nx, minS, s->lv.back().n);
  // fused path: grid-wide levels are even (restriction, tile parity), the tail from
  // the first level <= COARSE_N fits the one-workgroup kernel, and the sweep counts
  // have a compiled tile kernel
  int Lc = (int)s->lv.size();
  int cs = 0;
  while (cs < Lc && s->lv[cs].n > fsk::COARSE_N) cs++;
  s->fused = s->cfg.mg_fused && cs < Lc && Lc - cs <= fsk::CML_MAXL && s->cfg.mg_pre >= 1 && s->cfg.mg_pre <= 3 &&
             s->cfg.mg_post >= 1 && s->cfg.mg_post <= 3;
  s->bf16 = s->cfg.mg_precision == FS_MG_BF16;
  if (s->bf16 && !s->fused)
    return fail(FS_E_UNSUPPORTED, "BF16 multigrid storage needs the fused V-cycle: pre/post sweeps in 1..3 and at "
                                  "most %d levels at or below %d cells", fsk::CML_MAXL, fsk::COARSE_N);
  s->coarse_start = cs;
  for (auto& L : s->lv) L.tile[0] = L.tile[1] = tile_kind(L.n);
  for (size_t l = 0; l < s->lv.size(); l++) {
    size_t c = (size_t)s->lv[l].n * s->lv[l].n;
    size_t w = s->bf16 ? c / 2 : c;  // FP32 words (levels have an even cell count)
    int rc = dalloc(&s->lv[l].x, w);
    if (rc) return rc;
    if ((rc = dalloc(&s->lv[l].rhs, w))) return rc;
    // the corrections passed up are FP32 under every storage: rounding a smooth
    // correction to BF16 adds grid-scale noise whose Laplacian is as large as the
    // residual the V-cycle removed (at level 0 it goes straight into the FP64
    // pressure; at the levels below, prolongation carries it up)
    if (s->fused && (rc = dalloc(&s->lv[l].xb, c))) return rc;
  }
  if (s->cfg.mg_coarse_exact) return build_coarse_exact(s);
  return build_coarse_inverse(s);
}

// Time every tile shape of the down and up kernels on each grid-wide level (one
// warm-up and five timed launches on the zeroed buffers of a new solver) and keep
// the fastest. Level 0's down kernel is the fused residual + down-sweep; its
// per-tile partials go to a scratch row sized for the finest shape.
constexpr int kTuneRounds = 3, kTuneLaunches = 10;

// whether the V-cycle takes levels 1 and 0 up in one launch (enqueue_iteration's rule)
bool up2_level0(const fs_solver_t* s) {
  const int cs = s->coarse_start;
  if (!s->fused || cs < 2 || s->lv[0].n % 32) return false;
  const int pair = cs - 2 >= 1 && s->lv[cs - 1].n % 8 == 0 ? cs - 2 : cs;
  const int l = pair < cs && s->lv[pair].n % 16 == 0 ? pair - 1 : cs - 1;
  return l >= 1;
}

int autotune_tiles(fs_solver_t* s) {
  const int cs = s->coarse_start;
  if (!s->fused || cs == 0) return FS_OK;
  // the level-0 down-sweep is timed as the iterations after the first launch it, reading
  // the pressure and the previous correction; its block maxima go to row 1 of these partials
  double* tmp = nullptr;
  int rc = dalloc(&tmp, (size_t)s->rp_cols + tile_count(TILE8, s->lv[0].n));
  if (rc) return rc;
  double* saved = s->rpart;
  s->rpart = tmp;
  hipEvent_t a = nullptr, b = nullptr;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) rc = fail(FS_E_DEVICE, "hipEventCreate");
  const int pre = s->cfg.mg_pre, post = s->cfg.mg_post;
  const bool up2_0 = up2_level0(s);
  for (int l = 0; l < cs && !rc; l++) {
    Level& L = s->lv[l];
    for (int dir = 0; dir < 2 && !rc; dir++) {
      // level 0 goes up together with level 1 (k_mgt_up2) in 32 x 32, 64 x 32 or 64 x 64 tiles
      const bool pair0 = dir == 1 && l == 0 && up2_0;
      auto run = [&]() -> int {
        if (pair0) return DISPATCH_SW(post, launch_up2, s, 0);
        if (dir == 1) return DISPATCH_SW(post, launch_up, s, l);
        if (l == 0) return DISPATCH_SW(pre, launch_down0, s, 1, 0, false, false);
        return DISPATCH_SW(pre, launch_down, s, l);
      };
      // three rounds over the shapes (order rotated each round), ten launches each
      // after an untimed one; a shape's time is its best round
      for (int k = 0; k < GRID_TILE_KINDS; k++) L.tune_us[dir][k] = 0;
      for (int round = 0; round < kTuneRounds && !rc; round++)
        for (int j = 0; j < GRID_TILE_KINDS && !rc; j++) {
          const int k = (j + round) % GRID_TILE_KINDS;
          if (pair0 && k != TILE32 && !(k == TILE64SQ && up2_64sq_ok(s)) &&
              !(k == TILE64 && up2_64sq_ok(s) && !fold_coarse(s, 0)))
            continue;
          L.tile[dir] = k;
          if ((rc = run())) break;
          if (hipEventRecord(a, s->stream) != hipSuccess) rc = fail(FS_E_DEVICE, "hipEventRecord");
          for (int r = 0; r < kTuneLaunches && !rc; r++) rc = run();
          if (!rc && hipEventRecord(b, s->stream) != hipSuccess) rc = fail(FS_E_DEVICE, "hipEventRecord");
          float ms = 0;
          if (!rc && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess))
            rc = fail(FS_E_DEVICE, "autotune timing failed");
          const float us = ms * 1000.f / kTuneLaunches;
          if (!rc && (L.tune_us[dir][k] == 0 || us < L.tune_us[dir][k])) L.tune_us[dir][k] = us;
        }
      int best = TILE32;
      for (int k = 0; k < GRID_TILE_KINDS; k++)
        if (L.tune_us[dir][k] > 0 && L.tune_us[dir][k] < L.tune_us[dir][best]) best = k;
      L.tile[dir] = best;
    }
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  s->rpart = saved;
  dfree(tmp);
  return rc;
}

// residual partial rows for `iterations` + 2 refinement iterations
int alloc_rpart(fs_solver_t* s, int iterations) {
  int rows = iterations + 2;
  int cols = s->fused && s->coarse_start > 0 ? tile_count(s->lv[0].tile[0], s->nx) : grid_blocks(s->nx, s->ny);
  double* d = nullptr;
  double* h = nullptr;
  int rc = dalloc(&d, (size_t)rows * cols);
  if (rc) return rc;
  if ((rc = halloc(&h, (size_t)rows * cols))) {
    dfree(d);
    return rc;
  }
  int* flags = nullptr;
  if ((rc = dalloc(&flags, (size_t)rows))) {
    dfree(d);
    hfree(h);
    return rc;
  }
  dfree(s->rpart);
  hfree(s->rpart_host);
  dfree(s->gate_flag);
  s->rpart = d;
  s->rpart_host = h;
  s->rp_rows = rows;
  s->rp_cols = cols;
  s->gate_flag = flags;
  s->gate_len = rows;
  return FS_OK;
}

void destroy_graphs(fs_solver_t* s);

int ensure_rpart(fs_solver_t* s, int iterations) {
  if (s->cfg.pressure_solver != FS_P_MULTIGRID || iterations + 2 <= s->rp_rows) return FS_OK;
  int rc = set_device(s);
  if (rc) return rc;
  if ((rc = sync_stream(s))) return rc;
  destroy_graphs(s);  // captured bodies hold the old row and flag pointers
  return alloc_rpart(s, iterations);
}

void release_guess(fs_solver_t* s);

void release(fs_solver_t* s) {
  if (!s) return;
  (void)hipSetDevice(s->dev);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  destroy_graphs(s);
  for (auto& f : s->inflight) (void)hipEventDestroy(f.ev);
  for (auto e : s->ev_pool) (void)hipEventDestroy(e);
  dfree(s->gate_flag);
  dfree(s->modes_dev);
  dfree(s->p_hist);
  dfree(s->dye[0]);
  dfree(s->dye[1]);
  dfree(s->out_img);
  dfree(s->bloom[0]);
  dfree(s->bloom[1]);
  hfree(s->ring);
  release_guess(s);
  dfree(s->spec_Z);
  dfree(s->spec_part);
  dfree(s->spec_E);
  dfree(s->ftab_col);
  dfree(s->ftab_row);
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
    dfree(l.xb);
  }
  dfree(s->rpart);
  hfree(s->rpart_host);
  dfree(s->pinv);
  dfree(s->ex_q);
  dfree(s->mpart);
  dfree(s->coef);
  hfree(s->dt_ring);
  dfree(s->diag_part);
  dfree(s->diag_out);
  dfree(s->solid_part);
  dfree(s->solid_imp);
  hfree(s->host);
  for (auto& pr : s->prof.pending) {
    (void)hipEventDestroy(pr.a);
    (void)hipEventDestroy(pr.b);
  }
  for (auto e : s->prof.pool) (void)hipEventDestroy(e);
  if (s->rb_event) (void)hipEventDestroy(s->rb_event);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int run_diag(fs_solver_t* s, double out[fsk::DIAG_W]) {
  int c = s->cur;
  dim3 g = grid2(s->nx, s->ny);
  fsk::k_diag<<<g, BLK, 0, s->stream>>>(s->u[c], s->v[c], s->nx, s->ny, 1.0 / s->dx, 1.0 / s->dy, s->diag_part);
  LAUNCHCHK();
  fsk::k_diag_final<<<1, 256, 0, s->stream>>>(s->diag_part, (int)(g.x * g.y), s->diag_out);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(s->host, s->diag_out, sizeof(double) * fsk::DIAG_W, hipMemcpyDeviceToHost, s->stream));
  int rc = sync_stream(s);
  if (rc) return rc;
  for (int k = 0; k < fsk::DIAG_W; k++) out[k] = s->host[k];
  return FS_OK;
}

}  // namespace

// ======================================================================= ABI


namespace {

// ------------------------------------------------------------ cavity batches

struct fs_batch_t {
  fs_batch_config cfg{};
  int dev = 0;
  hipStream_t stream = nullptr;
  int nsim = 0;
  std::vector<int> ext;      // internal index -> caller's index
  std::vector<int> intl;     // caller's index -> internal index
  std::vector<fsk::CavSim> sims_h;
  fsk::CavSim* sims = nullptr;
  int* blk0 = nullptr;       // first block of each cavity, then the cavity of each block
  int nblocks = 0;
  long cells = 0;
  int max_lg = 0;
  float *u[2] = {}, *v[2] = {}, *fu = nullptr, *fv = nullptr;
  int cur = 0;
  double* p64[2] = {};
  int pidx = 0;
  bool have_p = false;
  // multigrid size classes s = 2^lg (lg from max_lg down to 2): grids of the first
  // count[lg] simulations, at base[lg] in the level buffers
  long base[13] = {}, count[13] = {};
  float *lx = nullptr, *lr = nullptr;  // pre-smoothed corrections, right-hand sides
  float* lxb = nullptr;      // post-smoothed corrections (the coarse solve's at 64 x 64)
  bool fused = false;        // LDS-tiled level kernels
  long* cbase = nullptr;     // per simulation: its 64 x 64 level
  float* pinv4 = nullptr;    // pseudo-inverse of the 4 x 4 Neumann Laplacian (h = 1/4)
  float* ex_tab = nullptr;   // exact solve of every cavity's 64 x 64 level (DCT-II basis)
  double *rpart = nullptr, *rpart_host = nullptr;
  int rp_rows = 0;
  double* mpart = nullptr;
  int last_pred = 2;
  int64_t steps = 0, total_vcycles = 0;
  int last_vcycles = 0;
  double last_residual = 0;
  // replayed step heads (graph mode): one executable graph per (predicted V-cycles,
  // pressure parity, zero start, cfl); the launches depend on nothing else
  struct Head {
    int c, a0;
    bool zero_p;
    double cfl;
    hipGraphExec_t exec;
  };
  std::vector<Head> heads;
  int64_t graph_captures = 0, graph_replays = 0;
  // level-kernel tile shape (TileKind) per size class 2^lg and direction (0 down, 1 up),
  // the autotuner's timings per shape (us, 0: not timed) and its sweep trials
  int tile[13][2] = {};
  float tune_us[13][2][GRID_TILE_KINDS] = {};
  struct SweepTrial {
    int pre, post;
    double ms_per_step;
    int vcycles;
  };
  std::vector<SweepTrial> sweep_trials;
  long off = 0, lvl_cells = 0;  // cells of the velocity and level buffers
};

int bsync(fs_batch_t* b) {
  HIPCHK(hipStreamSynchronize(b->stream));
  return FS_OK;
}

template <typename K, typename... Args>
int blaunch(fs_batch_t* b, K kern, dim3 g, dim3 t, Args... args) {
  kern<<<g, t, 0, b->stream>>>(args...);
  LAUNCHCHK();
  return FS_OK;
}

inline unsigned nblk(long items) { return (unsigned)((items + 255) / 256); }

// LDS-tiled Neumann level kernels of size class s over its cnt grids, tile shape TS
template <class TS>
int nmt_down_t(fs_batch_t* b, int sw, float* x, const float* r, int s, long cnt, float* cr) {
  const int tiles_x = s / TS::TX, tpg = tiles_x * (s / TS::TY);
  const dim3 g((unsigned)(cnt * tpg)), t(TS::NT);
  const float h2 = 1.f / ((float)s * s), ih2 = (float)((double)s * s);
  if (sw == 1) return blaunch(b, fsk::k_nmt_down<1, TS>, g, t, x, r, s, tiles_x, tpg, h2, ih2, cr);
  if (sw == 2) return blaunch(b, fsk::k_nmt_downq<2, TS>, g, t, x, r, s, tiles_x, tpg, h2, ih2, cr);
  return blaunch(b, fsk::k_nmt_down<3, TS>, g, t, x, r, s, tiles_x, tpg, h2, ih2, cr);
}
template <class TS>
int nmt_up_t(fs_batch_t* b, int sw, const float* x, float* xb, const float* r, const float* xc, int s, long cnt) {
  const int tiles_x = s / TS::TX, tpg = tiles_x * (s / TS::TY);
  const dim3 g((unsigned)(cnt * tpg)), t(TS::NT);
  const float h2 = 1.f / ((float)s * s);
  if (sw == 1) return blaunch(b, fsk::k_nmt_up<1, TS>, g, t, x, xb, r, xc, s, tiles_x, tpg, h2);
  if (sw == 2) return blaunch(b, fsk::k_nmt_upq<2, TS>, g, t, x, xb, r, xc, s, tiles_x, tpg, h2);
  return blaunch(b, fsk::k_nmt_up<3, TS>, g, t, x, xb, r, xc, s, tiles_x, tpg, h2);
}
int nmt_down(fs_batch_t* b, int kind, int sw, float* x, const float* r, int s, long cnt, float* cr) {
  return kind == TILE64     ? nmt_down_t<fsk::Tile64>(b, sw, x, r, s, cnt, cr)
       : kind == TILE32   ? nmt_down_t<fsk::Tile32>(b, sw, x, r, s, cnt, cr)
       : kind == TILE16   ? nmt_down_t<fsk::Tile16>(b, sw, x, r, s, cnt, cr)
       : kind == TILE64SQ ? nmt_down_t<fsk::Tile64sq>(b, sw, x, r, s, cnt, cr)
       : kind == TILE32W  ? nmt_down_t<fsk::Tile32w>(b, sw, x, r, s, cnt, cr)
       : kind == TILE16W  ? nmt_down_t<fsk::Tile16w>(b, sw, x, r, s, cnt, cr)
                          : nmt_down_t<fsk::Tile8>(b, sw, x, r, s, cnt, cr);
}
int nmt_up(fs_batch_t* b, int kind, int sw, const float* x, float* xb, const float* r, const float* xc, int s,
           long cnt) {
  return kind == TILE64     ? nmt_up_t<fsk::Tile64>(b, sw, x, xb, r, xc, s, cnt)
       : kind == TILE32   ? nmt_up_t<fsk::Tile32>(b, sw, x, xb, r, xc, s, cnt)
       : kind == TILE16   ? nmt_up_t<fsk::Tile16>(b, sw, x, xb, r, xc, s, cnt)
       : kind == TILE64SQ ? nmt_up_t<fsk::Tile64sq>(b, sw, x, xb, r, xc, s, cnt)
       : kind == TILE32W  ? nmt_up_t<fsk::Tile32w>(b, sw, x, xb, r, xc, s, cnt)
       : kind == TILE16W  ? nmt_up_t<fsk::Tile16w>(b, sw, x, xb, r, xc, s, cnt)
                          : nmt_up_t<fsk::Tile8>(b, sw, x, xb, r, xc, s, cnt);
}

// refinement iteration `it`: residual of p64[pin] (+ the finest corrections),
// then (unless residual_only) one V-cycle over every size class
int batch_iteration(fs_batch_t* b, int it, int pin, bool zero_p, bool residual_only) {
  int rc;
  const int c = b->cur;
  if ((rc = blaunch(b, fsk::k_cav_residual64, dim3((b->nblocks + fsk::CAV_WPB - 1) / fsk::CAV_WPB),
                    dim3(64 * fsk::CAV_WPB), (const float*)b->u[c],
                    (const float*)b->v[c], (const double*)(zero_p ? nullptr : b->p64[pin]),
                    (const float*)(it > 0 ? b->lxb : nullptr), b->p64[1 - pin], b->lr, (const fsk::CavSim*)b->sims,
                    (const int*)b->blk0, b->nsim, b->rpart + (size_t)it * b->nblocks, b->nblocks)))
    return rc;
  if (residual_only) return FS_OK;
  const int pre = b->cfg.mg_pre, post = b->cfg.mg_post;
  for (int lg = b->max_lg; lg > 6; lg--) {  // grid-wide levels above 64 x 64, finest first
    const long cnt = b->count[lg];
    const int s = 1 << lg;
    float* x = b->lx + b->base[lg];
    const float* r = b->lr + b->base[lg];
    const float h2 = 1.f / ((float)s * s);
    if (b->fused) {
      if ((rc = nmt_down(b, b->tile[lg][0], pre, x, r, s, cnt, b->lr + b->base[lg - 1]))) return rc;
      continue;
    }
    const long half = cnt * s * (s / 2);
    if (pre == 0) {
      fsk::k_fill_f32<<<nblk(cnt * s * s), 256, 0, b->stream>>>(x, (size_t)cnt * s * s, 0.f);
      LAUNCHCHK();
    }
    for (int sw = 0; sw < pre; sw++)
      for (int colour = 0; colour < 2; colour++)
        if ((rc = blaunch(b, fsk::k_nm_rb, dim3(nblk(half)), dim3(256), x, r, lg, cnt, h2, colour,
                          (sw == 0 && colour == 0) ? 1 : 0)))
          return rc;
    if ((rc = blaunch(b, fsk::k_nm_restrict, dim3(nblk(cnt * (s / 2) * (s / 2))), dim3(256), (const float*)x, r, lg,
                      cnt, (float)((double)s * s), b->lr + b->base[lg - 1])))
      return rc;
  }
  if (b->ex_tab) {
    // every cavity's 64 x 64 level solved exactly, one workgroup each
    if ((rc = blaunch(b, fsk::k_mg_coarse_exact<float>, dim3(b->nsim), dim3(fsk::EXT), b->lxb, (const float*)b->lr,
                      64, (const long*)b->cbase, (const uint4*)b->ex_tab, fsk::gate_open())))
      return rc;
  } else if ((rc = blaunch(b, fsk::k_nm_coarse, dim3(b->nsim), dim3(1024), b->lxb, (const float*)b->lr,
                           (const long*)b->cbase, 64, pre, post, (const float*)b->pinv4))) {
    return rc;
  }
  for (int lg = 7; lg <= b->max_lg; lg++) {
    const long cnt = b->count[lg];
    const int s = 1 << lg;
    const float* x = b->lx + b->base[lg];
    float* xb = b->lxb + b->base[lg];
    const float* r = b->lr + b->base[lg];
    const float* xc = b->lxb + b->base[lg - 1];
    const float h2 = 1.f / ((float)s * s);
    if (b->fused) {
      if ((rc = nmt_up(b, b->tile[lg][1], post, x, xb, r, xc, s, cnt))) return rc;
      continue;
    }
    const long half = cnt * s * (s / 2);
    if ((rc = blaunch(b, fsk::k_nm_prolong, dim3(nblk(cnt * s * s)), dim3(256), xc, x, xb, lg, cnt))) return rc;
    for (int sw = 0; sw < post; sw++)
      for (int colour = 0; colour < 2; colour++)
        if ((rc = blaunch(b, fsk::k_nm_rb, dim3(nblk(half)), dim3(256), xb, r, lg, cnt, h2, colour, 0))) return rc;
  }
  return FS_OK;
}

int batch_read_rmax(fs_batch_t* b, int rows, std::vector<double>& rmax) {
  HIPCHK(hipMemcpyAsync(b->rpart_host, b->rpart, sizeof(double) * (size_t)rows * b->nblocks, hipMemcpyDeviceToHost,
                        b->stream));
  int rc = bsync(b);
  if (rc) return rc;
  rmax.assign(rows, 0.0);
  for (int k = 0; k < rows; k++) {
    const double* row = b->rpart_host + (size_t)k * b->nblocks;
    double m = 0;
    for (int q = 0; q < b->nblocks; q++) m = std::max(m, row[q]);
    rmax[k] = m;
  }
  return FS_OK;
}

// The step up to the projection's first read-back: per-cavity dt, advection and
// diffusion (u* lands back in u[cur]), then refinement iterations 0 .. c-1 and the
// residual of iteration c. Launches only, so it can be captured as a graph.
int batch_enqueue_head(fs_batch_t* b, double cfl, int c, int a0, bool zero_p) {
  int rc;
  const fsk::CavSim* S = b->sims;
  if ((rc = blaunch(b, fsk::k_cav_dt, dim3(b->nsim), dim3(256), (const double*)b->mpart, b->sims, (const int*)b->blk0,
                    b->nsim, cfl, 1.0)))
    return rc;
  const int c0 = b->cur;
  if (b->cfg.advection == FS_ADV_MACCORMACK) {
    if ((rc = blaunch(b, fsk::k_cav_advect_sl, dim3(b->nblocks), dim3(256), (const float*)b->u[c0],
                      (const float*)b->v[c0], b->fu, b->fv, S, (const int*)b->blk0, b->nsim)))
      return rc;
    if ((rc = blaunch(b, fsk::k_cav_advect_mc, dim3(b->nblocks), dim3(256), (const float*)b->u[c0],
                      (const float*)b->v[c0], (const float*)b->fu, (const float*)b->fv, b->u[1 - c0], b->v[1 - c0], S,
                      (const int*)b->blk0, b->nsim)))
      return rc;
  } else {
    if ((rc = blaunch(b, fsk::k_cav_advect_sl, dim3(b->nblocks), dim3(256), (const float*)b->u[c0],
                      (const float*)b->v[c0], b->u[1 - c0], b->v[1 - c0], S, (const int*)b->blk0, b->nsim)))
      return rc;
  }
  if ((rc = blaunch(b, fsk::k_cav_diffuse_solid, dim3(b->nblocks), dim3(256), (const float*)b->u[1 - c0],
                    (const float*)b->v[1 - c0], b->u[c0], b->v[c0], S, (const int*)b->blk0, b->nsim,
                    (float)b->cfg.mask_width, b->cfg.coupling_passes)))
    return rc;
  for (int it = 0; it < c; it++)
    if ((rc = batch_iteration(b, it, (a0 + it) & 1, it == 0 && zero_p, false))) return rc;
  return batch_iteration(b, c, (a0 + c) & 1, c == 0 && zero_p, true);
}

// After the head: read the residual rows back; stop when every simulation's FP64
// residual max is within the tolerance, else single further V-cycles.
int batch_finish_projection(fs_batch_t* b, int c, int a0, bool zero_p, int* cycles_out) {
  const double tol = b->cfg.tolerance;
  const int maxc = b->cfg.max_vcycles;
  int rc;
  std::vector<double> rmax;
  for (;;) {
    if ((rc = batch_read_rmax(b, c + 1, rmax))) return rc;
    double last = rmax[c];
    if (!std::isfinite(last)) {
      b->pidx = (a0 + c + 1) & 1;
      b->last_residual = last;
      *cycles_out = c;
      return fail(FS_E_NONFINITE, "non-finite residual in a cavity projection (V-cycle %d)", c);
    }
    if (last <= tol || c >= maxc) {
      int need = c;
      for (int k = 0; k <= c; k++)
        if (rmax[k] <= tol) {
          need = k;
          break;
        }
      b->last_pred = need;
      b->pidx = (a0 + c + 1) & 1;
      b->last_residual = last;
      *cycles_out = c;
      if (last > tol)
        return fail(FS_E_NOT_CONVERGED, "cavity projection residual %.3g > %.3g after %d V-cycles", last, tol, c);
      return FS_OK;
    }
    if ((rc = batch_iteration(b, c, (a0 + c) & 1, c == 0 && zero_p, false))) return rc;
    c++;
    if ((rc = batch_iteration(b, c, (a0 + c) & 1, false, true))) return rc;
  }
}

// The head of a step as a replayed graph: captured on first use of its key.
int batch_launch_head_graph(fs_batch_t* b, double cfl, int c, int a0, bool zero_p) {
  for (const auto& h : b->heads)
    if (h.c == c && h.a0 == a0 && h.zero_p == zero_p && h.cfl == cfl) {
      HIPCHK(hipGraphLaunch(h.exec, b->stream));
      b->graph_replays++;
      return FS_OK;
    }
  if (b->heads.size() >= 64) {  // a changing cfl or prediction: start the cache afresh
    HIPCHK(hipStreamSynchronize(b->stream));
    for (auto& h : b->heads) (void)hipGraphExecDestroy(h.exec);
    b->heads.clear();
  }
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = batch_enqueue_head(b, cfl, c, a0, zero_p);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &graph);
  if (rc || e != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc ? rc : fail(FS_E_DEVICE, "capturing the step failed: %s", hipGetErrorString(e));
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(FS_E_DEVICE, "instantiating the step graph failed: %s", hipGetErrorString(e));
  b->heads.push_back({c, a0, zero_p, cfl, exec});
  b->graph_captures++;
  HIPCHK(hipGraphLaunch(exec, b->stream));
  b->graph_replays++;
  return FS_OK;
}

// The batch at rest: zero velocities, pressures and level buffers, step counters reset.
int batch_reset(fs_batch_t* b) {
  const long off = b->off, lvl = b->lvl_cells;
  for (float* a : {b->u[0], b->v[0], b->u[1], b->v[1], b->fu, b->fv})
    fsk::k_fill_f32<<<nblk(off), 256, 0, b->stream>>>(a, (size_t)off, 0.f);
  for (double* a : {b->p64[0], b->p64[1]}) fsk::k_fill_f64<<<nblk(off), 256, 0, b->stream>>>(a, (size_t)off, 0.0);
  for (float* a : {b->lx, b->lr, b->lxb}) fsk::k_fill_f32<<<nblk(lvl), 256, 0, b->stream>>>(a, (size_t)lvl, 0.f);
  fsk::k_fill_f64<<<nblk(2L * b->nblocks), 256, 0, b->stream>>>(b->mpart, (size_t)2 * b->nblocks, 0.0);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess)
    return fail(FS_E_DEVICE, "filling the batch state failed");
  b->cur = 0;
  b->pidx = 0;
  b->have_p = false;
  b->last_pred = 2;
  b->steps = b->total_vcycles = 0;
  b->last_vcycles = 0;
  b->last_residual = 0;
  return FS_OK;
}

// Tile shape of each size class's down and up kernels: every shape computes the
// same bits, so each is timed on the level buffers (best of three rounds of ten
// launches, the order rotated each round) and the fastest kept.
int batch_autotune_tiles(fs_batch_t* b) {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = FS_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = fail(FS_E_DEVICE, "hipEventCreate failed");
  for (int lg = b->max_lg; lg > 6 && !rc; lg--) {
    const long cnt = b->count[lg];
    const int s = 1 << lg;
    float* x = b->lx + b->base[lg];
    float* xb = b->lxb + b->base[lg];
    const float* r = b->lr + b->base[lg];
    float* cr = b->lr + b->base[lg - 1];
    const float* xc = b->lxb + b->base[lg - 1];
    for (int dir = 0; dir < 2 && !rc; dir++) {
      auto run = [&](int kind) {
        return dir == 0 ? nmt_down(b, kind, b->cfg.mg_pre, x, r, s, cnt, cr)
                        : nmt_up(b, kind, b->cfg.mg_post, x, xb, r, xc, s, cnt);
      };
      for (int round = 0; round < kTuneRounds && !rc; round++)
        for (int j = 0; j < GRID_TILE_KINDS && !rc; j++) {
          const int k = (j + round) % GRID_TILE_KINDS;
          if ((rc = run(k))) break;
          if (hipEventRecord(e0, b->stream) != hipSuccess) rc = fail(FS_E_DEVICE, "hipEventRecord failed");
          for (int q = 0; q < kTuneLaunches && !rc; q++) rc = run(k);
          if (!rc && hipEventRecord(e1, b->stream) != hipSuccess) rc = fail(FS_E_DEVICE, "hipEventRecord failed");
          float ms = 0;
          if (!rc && (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            rc = fail(FS_E_DEVICE, "autotune timing failed");
          if (rc) break;
          const float us = ms * 1000.f / kTuneLaunches;
          if (b->tune_us[lg][dir][k] == 0 || us < b->tune_us[lg][dir][k]) b->tune_us[lg][dir][k] = us;
        }
      for (int k = 0; k < GRID_TILE_KINDS; k++)
        if (b->tune_us[lg][dir][k] > 0 && b->tune_us[lg][dir][k] < b->tune_us[lg][dir][b->tile[lg][dir]])
          b->tile[lg][dir] = k;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (!rc) rc = batch_reset(b);
  return rc;
}

// Sweeps per V-cycle: each (pre, post) in {1, 2, 3}^2 runs the batch from rest for
// four steps at CFL 0.5 (launched one by one); the wall time of steps 2-4 decides.
// The choice changes the bits (every choice meets the same tolerance).
int batch_autotune_sweeps(fs_batch_t* b) {
  const int graph = b->cfg.graph;
  b->cfg.graph = 0;
  int rc = FS_OK;
  double best = 0;
  int bp = b->cfg.mg_pre, bq = b->cfg.mg_post;
  for (int pre = 1; pre <= 3 && !rc; pre++)
    for (int post = 1; post <= 3 && !rc; post++) {
      b->cfg.mg_pre = pre;
      b->cfg.mg_post = post;
      if ((rc = batch_reset(b)) || (rc = ::fs_batch_step_cfl(b, 0.5))) break;
      auto t0 = std::chrono::steady_clock::now();
      int64_t v0 = b->total_vcycles;
      for (int k = 0; k < 3 && !rc; k++) rc = ::fs_batch_step_cfl(b, 0.5);
      if (!rc) rc = bsync(b);
      if (rc) break;
      double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / 3.0;
      b->sweep_trials.push_back({pre, post, ms, (int)(b->total_vcycles - v0)});
      if (best == 0 || ms < best) {
        best = ms;
        bp = pre;
        bq = post;
      }
    }
  b->cfg.mg_pre = bp;
  b->cfg.mg_post = bq;
  b->cfg.graph = graph;
  if (!rc) rc = batch_reset(b);
  return rc;
}

void batch_release(fs_batch_t* b) {
  if (!b) return;
  (void)hipSetDevice(b->dev);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (auto& h : b->heads) (void)hipGraphExecDestroy(h.exec);
  b->heads.clear();
  for (int k = 0; k < 2; k++) {
    dfree(b->u[k]);
    dfree(b->v[k]);
    dfree(b->p64[k]);
  }
  dfree(b->fu);
  dfree(b->fv);
  dfree(b->sims);
  dfree(b->blk0);
  dfree(b->lx);
  dfree(b->lr);
  dfree(b->lxb);
  dfree(b->cbase);
  dfree(b->pinv4);
  dfree(b->ex_tab);
  dfree(b->rpart);
  hfree(b->rpart_host);
  dfree(b->mpart);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// pseudo-inverse of the 4 x 4 Laplacian with zero-gradient walls, h = 1/4, transposed:
// A+ = (A + 11^T/N)^-1 - 11^T/N (null space: the constants)
std::vector<float> neumann_pinv4() {
  const int n = 4, N = 16;
  const double h2 = 1.0 / 16.0;
  std::vector<double> M((size_t)N * 2 * N, 0.0);
  auto at = [&](int r, int c) -> double& { return M[(size_t)r * 2 * N + c]; };
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      int r = j * n + i;
      int nb[4][2] = {{i - 1, j}, {i + 1, j}, {i, j - 1}, {i, j + 1}};
      for (auto& q : nb)
        if (q[0] >= 0 && q[0] < n && q[1] >= 0 && q[1] < n) {
          at(r, q[1] * n + q[0]) += 1.0 / h2;
          at(r, r) -= 1.0 / h2;
        }
      for (int c = 0; c < N; c++) at(r, c) += 1.0 / N;
      at(r, N + r) = 1.0;
    }
  for (int col = 0; col < N; col++) {
    int piv = col;
    for (int r = col + 1; r < N; r++)
      if (std::fabs(at(r, col)) > std::fabs(at(piv, col))) piv = r;
    if (piv != col)
      for (int c = 0; c < 2 * N; c++) std::swap(at(piv, c), at(col, c));
    double d = at(col, col);
    for (int c = 0; c < 2 * N; c++) at(col, c) /= d;
    for (int r = 0; r < N; r++)
      if (r != col) {
        double f = at(r, col);
        if (f != 0.0)
          for (int c = 0; c < 2 * N; c++) at(r, c) -= f * at(col, c);
      }
  }
  std::vector<float> pt((size_t)N * N);
  for (int r = 0; r < N; r++)
    for (int c = 0; c < N; c++) pt[(size_t)c * N + r] = (float)(at(r, N + c) - 1.0 / N);
  return pt;
}

}  // namespace

// ------------------------------------------------------------ guess trainer

namespace {

// Training state of the learned guess for batches of B images at n x n.
struct fs_trainer_t {
  int dev = 0;
  hipStream_t stream = nullptr;
  int n = 0, B = 0;
  double h = 0;
  size_t P = 0;                              // parameters, blob order
  size_t woff[15] = {}, boff[15] = {};       // offsets of each layer's weights and bias
  float *par = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
  int64_t adam_steps = 0;
  bool loaded = false, have_batch = false;
  uint16_t *wf[15] = {}, *wt[15] = {};       // BF16 weights: forward, transposed and tap-flipped
  uint16_t *act[14] = {}, *gz[14] = {};      // activations and pre-activation gradients, NHWC BF16
  float* gcat[3] = {};                       // gradients reaching enc0b, enc1b, enc2b through the skips
  double* pp = nullptr;                      // p_prev of the batch (FP64)
  float *r32 = nullptr, *tgt = nullptr;      // residual (FP32) and target
  float* err = nullptr;                      // output minus target
  double *scales = nullptr, *mr = nullptr, *mp = nullptr;
  int pblk = 0;                              // blocks per image of the pair preparation
  float* part = nullptr;                     // per-workgroup gradient partials
  float* part2 = nullptr;                    // their sums over groups of TR_RC rows
  size_t part_cap = 0;
  float* opart = nullptr;
  double* lpart = nullptr;
  int oblk = 0;
  double *loss_d = nullptr, *loss_h = nullptr;
  int *idx_d = nullptr, *idx_h = nullptr;
  double* snap = nullptr;                    // pressure snapshots, FP64 [nsnap][n][n]
  size_t nsnap = 0;
};

void trainer_release(fs_trainer_t* t) {
  if (!t) return;
  if (t->stream) {
    (void)hipSetDevice(t->dev);
    (void)hipStreamSynchronize(t->stream);
  }
  for (void* q : {(void*)t->par, (void*)t->grad, (void*)t->m1, (void*)t->m2, (void*)t->pp, (void*)t->r32,
                  (void*)t->tgt, (void*)t->err, (void*)t->scales, (void*)t->mr, (void*)t->mp, (void*)t->part, (void*)t->opart,
                  (void*)t->lpart, (void*)t->loss_d, (void*)t->idx_d, (void*)t->snap,
                  (void*)t->part2})
    dfree(q);
  for (int l = 0; l < 15; l++) {
    dfree(t->wf[l]);
    dfree(t->wt[l]);
  }
  for (int a = 0; a < 14; a++) {
    dfree(t->act[a]);
    dfree(t->gz[a]);
  }
  for (int a = 0; a < 3; a++) dfree(t->gcat[a]);
  hfree(t->loss_h);
  hfree(t->idx_h);
  if (t->stream) (void)hipStreamDestroy(t->stream);
  delete t;
}

template <typename K, typename... Args>
int tlaunch(fs_trainer_t* t, K kern, dim3 g, dim3 b, Args... args) {
  kern<<<g, b, 0, t->stream>>>(args...);
  LAUNCHCHK();
  return FS_OK;
}

// dst[i] = sum over r < rows of part[r][i] (len columns): in groups of TR_RC rows, then
// over the groups; the same fixed order every time.
int train_reduce(fs_trainer_t* t, const float* part, int rows, size_t len, float* dst) {
  const unsigned gx = (unsigned)((len + 255) / 256);
  if (rows <= fsk::TR_RC) return tlaunch(t, fsk::k_tr_reduce, dim3(gx), dim3(256), part, rows, len, dst);
  const int groups = (rows + fsk::TR_RC - 1) / fsk::TR_RC;
  int rc = tlaunch(t, fsk::k_tr_reduce_rows, dim3(gx, (unsigned)groups), dim3(256), part, rows, len, t->part2);
  if (rc) return rc;
  return tlaunch(t, fsk::k_tr_reduce, dim3(gx), dim3(256), (const float*)t->part2, groups, len, dst);
}

template <int CIN, int COUT, int MODE, int C2, int TH, int NBW, int WC = 1>
int train_fwd(fs_trainer_t* t, int layer) {
  const int m = t->n >> kGuessLayers[layer].level;
  const uint16_t* x1 = t->act[kGuessIn[layer][0]];
  const uint16_t* x2 = kGuessIn[layer][1] >= 0 ? t->act[kGuessIn[layer][1]] : nullptr;
  const int ntiles = (m / fsk::GT_W) * (m / TH) * t->B, nyb = COUT / (16 * NBW);
  dim3 g((unsigned)std::max(1, std::min(ntiles, 2048 / nyb)), (unsigned)nyb);
  return tlaunch(t, fsk::k_conv3x3<CIN, COUT, MODE, C2, TH, NBW, fsk::EPI_RELU, 0, WC>, g, dim3(fsk::GT_NT * WC), x1, x2,
                 (const uint16_t*)t->wf[layer], (const float*)(t->par + t->boff[layer]), t->act[layer], m, t->B,
                 fsk::ConvEpiArgs{});
}

// Data gradient of 3x3 layer `layer` (CIN, COUT its own channels): k_conv3x3 over
// gz[layer] with the transposed weights, COUT -> CIN channels, epilogue EPI.
template <int CIN, int COUT, int TH, int NBW, int EPI, int OS>
int train_bwd_data(fs_trainer_t* t, int layer, uint16_t* y, fsk::ConvEpiArgs e) {
  const int m = t->n >> kGuessLayers[layer].level;
  const int ntiles = (m / fsk::GT_W) * (m / TH) * t->B, nyb = CIN / (16 * NBW);
  dim3 g((unsigned)std::max(1, std::min(ntiles, 2048 / nyb)), (unsigned)nyb);
  return tlaunch(t, fsk::k_conv3x3<COUT, CIN, fsk::CONV_DIRECT, 0, TH, NBW, EPI, OS>, g, dim3(fsk::GT_NT),
                 (const uint16_t*)t->gz[layer], (const uint16_t*)nullptr, (const uint16_t*)t->wt[layer],
                 (const float*)nullptr, y, m, t->B, e);
}

// Weight and bias gradient of 3x3 layer `layer` into grad (partials, then a fixed-order sum).
template <int CIN, int COUT, int MODE, int C2, int TH>
int train_bwd_weights(fs_trainer_t* t, int layer, bool size_only = false, size_t* need = nullptr) {
  const int m = t->n >> kGuessLayers[layer].level;
  const int ntiles = (m / fsk::GT_W) * (m / TH) * t->B, nyb = COUT / 16;
  const int gx = std::max(1, std::min(std::max(1, ntiles / 4), 2048 / nyb));
  const size_t PL = (size_t)COUT * 9 * CIN + COUT;
  if (size_only) {
    *need = std::max(*need, PL * (size_t)gx);
    return FS_OK;
  }
  const uint16_t* x1 = t->act[kGuessIn[layer][0]];
  const uint16_t* x2 = kGuessIn[layer][1] >= 0 ? t->act[kGuessIn[layer][1]] : nullptr;
  int rc;
  if ((rc = tlaunch(t, fsk::k_conv3x3_wgrad<CIN, COUT, MODE, C2, TH>, dim3((unsigned)gx, (unsigned)nyb),
                    dim3(fsk::GT_NT), x1, x2, (const uint16_t*)t->gz[layer], t->part, m, t->B)))
    return rc;
  return train_reduce(t, t->part, gx, PL, t->grad + t->woff[layer]);
}

// every weight-gradient launch (or only the partial-buffer size it needs)
int train_all_weights(fs_trainer_t* t, bool size_only, size_t* need) {
  using namespace fsk;
  int rc;
  if ((rc = train_bwd_weights<16, 16, CONV_DIRECT, 0, 16>(t, 13, size_only, need))) return rc;
  if ((rc = train_bwd_weights<48, 16, CONV_UPCAT, 32, 16>(t, 12, size_only, need))) return rc;
  if ((rc = train_bwd_weights<32, 32, CONV_DIRECT, 0, 16>(t, 11, size_only, need))) return rc;
  if ((rc = train_bwd_weights<96, 32, CONV_UPCAT, 64, 16>(t, 10, size_only, need))) return rc;
  if ((rc = train_bwd_weights<64, 64, CONV_DIRECT, 0, 8>(t, 9, size_only, need))) return rc;
  if ((rc = train_bwd_weights<128, 64, CONV_UPCAT, 64, 8>(t, 8, size_only, need))) return rc;
  if ((rc = train_bwd_weights<64, 64, CONV_DIRECT, 0, 8>(t, 7, size_only, need))) return rc;
  if ((rc = train_bwd_weights<64, 64, CONV_POOL, 0, 8>(t, 6, size_only, need))) return rc;
  if ((rc = train_bwd_weights<64, 64, CONV_DIRECT, 0, 8>(t, 5, size_only, need))) return rc;
  if ((rc = train_bwd_weights<32, 64, CONV_POOL, 0, 8>(t, 4, size_only, need))) return rc;
  if ((rc = train_bwd_weights<32, 32, CONV_DIRECT, 0, 16>(t, 3, size_only, need))) return rc;
  if ((rc = train_bwd_weights<16, 32, CONV_POOL, 0, 16>(t, 2, size_only, need))) return rc;
  if ((rc = train_bwd_weights<16, 16, CONV_DIRECT, 0, 16>(t, 1, size_only, need))) return rc;
  return FS_OK;
}

int train_first_grid(const fs_trainer_t* t) {
  const int ntiles = (t->n / 16) * (t->n / 16) * t->B;
  return std::max(1, std::min(ntiles / 4, 1024));
}

// Forward pass, loss and every gradient of the batch whose indices are in idx_d.
int train_forward_backward(fs_trainer_t* t) {
  using namespace fsk;
  const int n = t->n, B = t->B;
  const size_t nn = (size_t)n * n;
  const double h2 = t->h * t->h;
  int rc;
  // ---- pairs, scales, targets
  if ((rc = tlaunch(t, k_tr_prep, dim3((unsigned)t->pblk, (unsigned)B), dim3(256), (const double*)t->snap,
                    (const int*)t->idx_d, n, 1.0 / h2, t->pp, t->r32, t->mr, t->mp)))
    return rc;
  if ((rc = tlaunch(t, k_tr_scales, dim3((unsigned)B), dim3(256), (const double*)t->mr, (const double*)t->mp, t->pblk,
                    t->scales)))
    return rc;
  if ((rc = tlaunch(t, k_tr_target, dim3((unsigned)t->pblk, (unsigned)B), dim3(256), (const double*)t->snap,
                    (const int*)t->idx_d, n, h2, (const double*)t->scales, t->tgt)))
    return rc;
  // ---- forward (the inference kernels over the batch)
  if ((rc = tlaunch(t, k_guess_first<float>, dim3((unsigned)(n / 64), (unsigned)(n / 4), (unsigned)B), dim3(256),
                    (const float*)t->r32, (const double*)t->pp, (const double*)t->scales,
                    (const float*)(t->par + t->woff[0]), (const float*)(t->par + t->boff[0]), t->act[0], n)))
    return rc;
  if ((rc = train_fwd<16, 16, CONV_DIRECT, 0, 8, 1>(t, 1))) return rc;
  if ((rc = train_fwd<16, 32, CONV_POOL, 0, 8, 2, 2>(t, 2))) return rc;
  if ((rc = train_fwd<32, 32, CONV_DIRECT, 0, 16, 2>(t, 3))) return rc;
  if ((rc = train_fwd<32, 64, CONV_POOL, 0, 8, 4, 2>(t, 4))) return rc;
  if ((rc = train_fwd<64, 64, CONV_DIRECT, 0, 16, 4, 2>(t, 5))) return rc;
  if ((rc = train_fwd<64, 64, CONV_POOL, 0, 8, 2, 2>(t, 6))) return rc;
  if ((rc = train_fwd<64, 64, CONV_DIRECT, 0, 16, 1>(t, 7))) return rc;
  if ((rc = train_fwd<128, 64, CONV_UPCAT, 64, 16, 2, 2>(t, 8))) return rc;
  if ((rc = train_fwd<64, 64, CONV_DIRECT, 0, 16, 4, 2>(t, 9))) return rc;
  if ((rc = train_fwd<96, 32, CONV_UPCAT, 64, 16, 2, 2>(t, 10))) return rc;
  if ((rc = train_fwd<32, 32, CONV_DIRECT, 0, 16, 2>(t, 11))) return rc;
  if ((rc = train_fwd<48, 16, CONV_UPCAT, 32, 16, 1>(t, 12))) return rc;
  if ((rc = train_fwd<16, 16, CONV_DIRECT, 0, 8, 1>(t, 13))) return rc;
  // ---- output layer: loss, its gradients, gz of dec0b
  const size_t count = nn * B;
  if ((rc = tlaunch(t, k_tr_out, dim3((unsigned)t->oblk), dim3(256), (const uint16_t*)t->act[13],
                    (const float*)(t->par + t->woff[14]), (const float*)(t->par + t->boff[14]), (const float*)t->tgt,
                    count, t->gz[13], t->lpart, t->opart, t->err)))
    return rc;
  if ((rc = tlaunch(t, k_tr_loss, dim3(1), dim3(256), (const double*)t->lpart, t->oblk, (double)count, t->loss_d)))
    return rc;
  if ((rc = train_reduce(t, t->opart, t->oblk, 17, t->grad + t->woff[14]))) return rc;
  // ---- data gradients, deepest first (each skip gradient is written before it is added)
  auto E = [](const uint16_t* act, const float* gin, float* gout) { return ConvEpiArgs{act, gin, gout}; };
  if ((rc = train_bwd_data<16, 16, 16, 1, EPI_MASK, 0>(t, 13, t->gz[12], E(t->act[12], nullptr, nullptr)))) return rc;
  if ((rc = train_bwd_data<48, 16, 16, 1, EPI_UPCAT, 32>(t, 12, t->gz[11], E(t->act[11], nullptr, t->gcat[0]))))
    return rc;
  if ((rc = train_bwd_data<32, 32, 16, 2, EPI_MASK, 0>(t, 11, t->gz[10], E(t->act[10], nullptr, nullptr)))) return rc;
  if ((rc = train_bwd_data<96, 32, 16, 2, EPI_UPCAT, 64>(t, 10, t->gz[9], E(t->act[9], nullptr, t->gcat[1]))))
    return rc;
  if ((rc = train_bwd_data<64, 64, 8, 2, EPI_MASK, 0>(t, 9, t->gz[8], E(t->act[8], nullptr, nullptr)))) return rc;
  if ((rc = train_bwd_data<128, 64, 8, 2, EPI_UPCAT, 64>(t, 8, t->gz[7], E(t->act[7], nullptr, t->gcat[2]))))
    return rc;
  if ((rc = train_bwd_data<64, 64, 8, 1, EPI_MASK, 0>(t, 7, t->gz[6], E(t->act[6], nullptr, nullptr)))) return rc;
  if ((rc = train_bwd_data<64, 64, 8, 1, EPI_UNPOOL, 0>(t, 6, t->gz[5], E(t->act[5], t->gcat[2], nullptr))))
    return rc;
  if ((rc = train_bwd_data<64, 64, 8, 2, EPI_MASK, 0>(t, 5, t->gz[4], E(t->act[4], nullptr, nullptr)))) return rc;
  if ((rc = train_bwd_data<32, 64, 8, 2, EPI_UNPOOL, 0>(t, 4, t->gz[3], E(t->act[3], t->gcat[1], nullptr))))
    return rc;
  if ((rc = train_bwd_data<32, 32, 16, 2, EPI_MASK, 0>(t, 3, t->gz[2], E(t->act[2], nullptr, nullptr)))) return rc;
  if ((rc = train_bwd_data<16, 32, 16, 1, EPI_UNPOOL, 0>(t, 2, t->gz[1], E(t->act[1], t->gcat[0], nullptr))))
    return rc;
  if ((rc = train_bwd_data<16, 16, 16, 1, EPI_MASK, 0>(t, 1, t->gz[0], E(t->act[0], nullptr, nullptr)))) return rc;
  // ---- weight gradients
  if ((rc = train_all_weights(t, false, nullptr))) return rc;
  const int g0 = train_first_grid(t);
  if ((rc = tlaunch(t, k_tr_first_wgrad, dim3((unsigned)g0), dim3(320), (const uint16_t*)t->gz[0],
                    (const float*)t->r32, (const double*)t->pp, (const double*)t->scales, n, B, t->part)))
    return rc;
  return train_reduce(t, t->part, g0, 304, t->grad + t->woff[0]);
}

int train_pack(fs_trainer_t* t) {
  for (int l = 1; l < 14; l++) {
    const GuessLayer& L = kGuessLayers[l];
    int nw = L.cout * 9 * L.cin;
    int rc = tlaunch(t, fsk::k_tr_pack, dim3((unsigned)((nw + 255) / 256)), dim3(256),
                     (const float*)(t->par + t->woff[l]), L.cin, L.cout, t->wf[l], t->wt[l]);
    if (rc) return rc;
  }
  return FS_OK;
}

}  // namespace

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
  c->autotune = 1;
  c->adv_fused = 1;
  c->device = 0;
  c->mg_fused = 1;
  c->gated = 1;
  c->graph = 0;
  c->dye_fade = 0.0;
  c->mg_coarse_exact = 0;
  c->tile = 0;
  c->confinement = 0.0;
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
  if ((c.pressure_solver == FS_P_RED_BLACK || c.pressure_solver == FS_P_MULTIGRID) && ((c.nx | c.ny) & 1))
    return fail(FS_E_INVALID_ARG,
                "red-black ordering needs even grid dimensions on a periodic grid, got %dx%d", c.nx, c.ny);
  if (c.tolerance < 0 || !std::isfinite(c.tolerance)) return fail(FS_E_INVALID_ARG, "tolerance must be >= 0");
  if (c.tolerance > 0 && c.pressure_solver != FS_P_MULTIGRID)
    return fail(FS_E_INVALID_ARG, "a residual tolerance needs the multigrid solver");
  if (c.mg_pre < 0 || c.mg_post < 0 || c.mg_coarse < 0) return fail(FS_E_INVALID_ARG, "negative sweep count");
  if (!std::isfinite(c.confinement)) return fail(FS_E_INVALID_ARG, "confinement must be finite");
  if (c.mg_coarse_exact != 0 && c.mg_coarse_exact != 1)
    return fail(FS_E_INVALID_ARG, "mg_coarse_exact must be 0 or 1, got %d", c.mg_coarse_exact);
  if (c.tile < 0 || c.tile > GRID_TILE_KINDS)
    return fail(FS_E_INVALID_ARG, "tile must be in 0..%d, got %d", GRID_TILE_KINDS, c.tile);
  if (c.mg_precision == FS_MG_FP16)
    return fail(FS_E_UNSUPPORTED,
                "FP16 multigrid storage is not built: the correction scales as h^2 |r| and underflows FP16 on "
                "fine grids; use BF16 (FP32 exponent range)");
  if (c.mg_precision != FS_MG_FP32 && c.mg_precision != FS_MG_BF16)
    return fail(FS_E_INVALID_ARG, "unknown multigrid precision %d", c.mg_precision);
  if (c.mg_precision == FS_MG_BF16 && !(c.pressure_solver == FS_P_MULTIGRID && c.mg_fused))
    return fail(FS_E_UNSUPPORTED, "BF16 multigrid storage needs the fused multigrid V-cycle (mg_fused = 1)");
  if (c.guess == FS_GUESS_LEARNED || c.guess == FS_GUESS_LEARNED_COARSE || c.guess == FS_GUESS_LEARNED_QUARTER)
    return fail(FS_E_INVALID_ARG, "create with guess zero or previous, then fs_load_guess and fs_set_guess(learned)");
  if (c.guess != FS_GUESS_ZERO && c.guess != FS_GUESS_PREVIOUS && c.guess != FS_GUESS_EXTRAPOLATED)
    return fail(FS_E_INVALID_ARG, "unknown guess mode %d", c.guess);
  if (c.guess == FS_GUESS_EXTRAPOLATED && c.pressure_solver != FS_P_MULTIGRID)
    return fail(FS_E_UNSUPPORTED, "the extrapolated guess needs the multigrid solver");
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
  s->gated = c.gated != 0;
  s->use_graph = c.graph != 0;
  auto bail = [&](int code) {
    std::string msg = g_err;
    release(s);
    g_err = msg;
    return code;
  };
  if (hipSetDevice(s->dev) != hipSuccess) return bail(fail(FS_E_DEVICE, "hipSetDevice(%d) failed", s->dev));
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FS_E_DEVICE, "hipStreamCreate failed"));
  if (hipEventCreateWithFlags(&s->rb_event, hipEventDisableTiming) != hipSuccess)
    return bail(fail(FS_E_DEVICE, "hipEventCreate failed"));
  int rc = FS_OK;
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
  if (!rc && c.guess == FS_GUESS_EXTRAPOLATED) rc = dalloc(&s->p_hist, s->n);
  if (!rc) rc = build_levels(s);
  s->mp_count = grid_blocks(s->nx, s->ny);
  if (!rc) rc = dalloc(&s->mpart, (size_t)2 * s->mp_count);
  if (!rc) rc = dalloc(&s->coef, 1);
  if (!rc && hipMemsetAsync(s->coef, 0, sizeof(fsk::StepCoef), s->stream) != hipSuccess)
    rc = fail(FS_E_DEVICE, "clearing the step coefficients failed");
  if (!rc) rc = dalloc(&s->modes_dev, 1);
  if (!rc) rc = halloc_coherent(&s->ring, REC_RING);
  if (!rc) rc = halloc(&s->dt_ring, DT_RING);
  if (!rc) rc = dalloc(&s->diag_part, (size_t)grid_blocks(s->nx, s->ny) * fsk::DIAG_W);
  if (!rc) rc = dalloc(&s->diag_out, fsk::DIAG_W);
  if (!rc) rc = dalloc(&s->solid_part, (size_t)grid_blocks(s->nx, s->ny) * 3);
  if (!rc) rc = dalloc(&s->solid_imp, 3);
  if (!rc) rc = halloc(&s->host, 64);
  if (rc) return bail(rc);
  // zero state: velocity, pressure, multigrid buffers
  if (!rc) rc = fill_f32(s, s->u[0], s->n, 0.f);
  if (!rc) rc = fill_f32(s, s->v[0], s->n, 0.f);
  if (!rc) rc = fill_f64(s, s->p64[0], s->n, 0.0);
  if (!rc) rc = fill_f64(s, s->p64[1], s->n, 0.0);
  if (!rc) rc = fill_f32(s, s->div, s->n, 0.f);
  for (auto& l : s->lv) {  // zero bits are 0.0 in FP32 and in BF16
    size_t w = s->bf16 ? (size_t)l.n * l.n / 2 : (size_t)l.n * l.n;
    if (!rc) rc = fill_f32(s, l.x, w, 0.f);
    if (!rc) rc = fill_f32(s, l.rhs, w, 0.f);
    if (!rc && l.xb) rc = fill_f32(s, l.xb, (size_t)l.n * l.n, 0.f);
  }
  // on zero inputs every fused kernel writes zeros, so tuning leaves the state as it is
  if (!rc && c.pressure_solver == FS_P_MULTIGRID && c.tile > 0) {
    // one shape for every grid-wide level (the levels-1-and-0 launch takes 32 x 32 unless
    // the shape is 64 x 64 and the grid is whole 64 x 64 tiles)
    for (int l = 0; l < s->coarse_start; l++) s->lv[l].tile[0] = s->lv[l].tile[1] = c.tile - 1;
  } else if (!rc && c.pressure_solver == FS_P_MULTIGRID && c.autotune) {
    rc = autotune_tiles(s);
  }
  if (!rc && c.pressure_solver == FS_P_MULTIGRID) rc = alloc_rpart(s, std::max(c.pressure_iterations, 1));
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
  if ((rc = sync_stream(s))) return rc;
  s->have_maxabs = false;
  s->have_p = false;
  s->have_hist = false;
  s->pred_known = false;
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
  return sync_stream(s);
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
    case FS_F_DYE:
      if (!s->dye[0]) return fail(FS_E_INVALID_ARG, "no dye: upload one or splat some first");
      src = s->dye[s->dye_cur];
      need = s->n * 4 * sizeof(float);
      break;
    case FS_F_OUTPUT:
      if (!s->out_img) return fail(FS_E_INVALID_ARG, "nothing rendered yet");
      src = s->out_img;
      need = s->n * sizeof(fsk::Half4);
      break;
    case FS_F_BLOOM:
      if (!s->out_img) return fail(FS_E_INVALID_ARG, "nothing rendered yet");
      src = s->bloom[s->bloom_cur];
      need = (size_t)(s->nx >> 1) * (size_t)(s->ny >> 1) * sizeof(fsk::Half4);
      break;
    case FS_F_PRESSURE:
      need = s->n * sizeof(double);
      if (s->cfg.pressure_solver != FS_P_MULTIGRID) {
        if (bytes != need) return fail(FS_E_INVALID_ARG, "field %d needs %zu bytes, got %zu", field, need, bytes);
        if ((rc = set_device(s))) return rc;
        std::vector<float> pf(s->n);
        HIPCHK(hipMemcpyAsync(pf.data(), s->p32[s->p32idx], s->n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
        if ((rc = sync_stream(s))) return rc;
        for (size_t k = 0; k < s->n; k++) ((double*)out)[k] = pf[k];
        return FS_OK;
      }
      src = s->p64[s->pidx];
      break;
    case FS_F_RESIDUAL:
      if (s->lv.empty()) return fail(FS_E_INVALID_ARG, "no multigrid residual without the multigrid solver");
      src = s->lv[0].rhs;
      need = s->n * sizeof(float);
      if (s->bf16) {  // widen the stored BF16 values
        if (bytes != need) return fail(FS_E_INVALID_ARG, "field %d needs %zu bytes, got %zu", field, need, bytes);
        if ((rc = set_device(s))) return rc;
        std::vector<uint16_t> b(s->n);
        HIPCHK(hipMemcpyAsync(b.data(), src, s->n * sizeof(uint16_t), hipMemcpyDeviceToHost, s->stream));
        if ((rc = sync_stream(s))) return rc;
        for (size_t k = 0; k < s->n; k++) {
          uint32_t w = (uint32_t)b[k] << 16;
          std::memcpy((float*)out + k, &w, 4);
        }
        return FS_OK;
      }
      break;
    default: return fail(FS_E_INVALID_ARG, "unknown field %d", field);
  }
  if (bytes != need) return fail(FS_E_INVALID_ARG, "field %d needs %zu bytes, got %zu", field, need, bytes);
  if ((rc = set_device(s))) return rc;
  HIPCHK(hipMemcpyAsync(out, src, need, hipMemcpyDeviceToHost, s->stream));
  return sync_stream(s);
}

int fs_step(void* h, double dt) { return fs_step_splats(h, dt, nullptr, 0); }

}  // extern "C"

namespace {
// Allocates the dye (zero) on first use; captured step bodies then need it too.
int ensure_dye(fs_solver_t* s) {
  if (s->dye[0]) return FS_OK;
  int rc;
  if ((rc = sync_stream(s))) return rc;
  destroy_graphs(s);
  if ((rc = dalloc(&s->dye[0], s->n)) || (rc = dalloc(&s->dye[1], s->n))) return rc;
  if ((rc = fill_f32(s, (float*)s->dye[0], 4 * s->n, 0.f))) return rc;
  s->dye_cur = 0;
  return FS_OK;
}
}  // namespace

extern "C" {

int fs_step_splats(void* h, double dt, const fs_splat* splats, int count) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!(dt >= 0) || !std::isfinite(dt)) return fail(FS_E_INVALID_ARG, "dt must be finite and >= 0, got %g", dt);
  if (count < 0 || (count > 0 && !splats)) return fail(FS_E_INVALID_ARG, "splats is NULL or count < 0");
  if ((rc = set_device(s))) return rc;
  bool dye_splat = false;
  for (int m = 0; m < count; m++) {
    const fs_splat& p = splats[m];
    if (!std::isfinite(p.x) || !std::isfinite(p.y) || !std::isfinite(p.dx) || !std::isfinite(p.dy) ||
        !std::isfinite(p.r) || !std::isfinite(p.g) || !std::isfinite(p.b) || !std::isfinite(p.radius))
      return fail(FS_E_INVALID_ARG, "splat %d is not finite", m);
    dye_splat = dye_splat || p.is_dye;
  }
  if (dye_splat && (rc = ensure_dye(s))) return rc;
  s->splats.clear();
  for (int m = 0; m < count; m++) {
    const fs_splat& p = splats[m];
    s->splats.push_back(fsk::Splat{(float)p.x, (float)p.y, (float)p.dx, (float)p.dy, (float)p.r, (float)p.g,
                                   (float)p.b, (float)p.radius, p.is_dye ? 1 : 0});
  }
  if (gated_eligible(s)) {
    rc = step_gated(s, false, dt);
  } else if (!(rc = consume_records(s, true)) && !(rc = enqueue_head(s, false, dt, 0))) {
    rc = do_step(s);
  }
  s->splats.clear();
  return rc;
}

int fs_upload_dye(void* h, const float* rgba, size_t n) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!rgba) return fail(FS_E_INVALID_ARG, "rgba is NULL");
  if (n != 4 * s->n) return fail(FS_E_INVALID_ARG, "expected %zu values (rgba per cell), got %zu", 4 * s->n, n);
  if ((rc = set_device(s)) || (rc = ensure_dye(s))) return rc;
  HIPCHK(hipMemcpyAsync(s->dye[s->dye_cur], rgba, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
  return sync_stream(s);
}

int fs_render(void* h, double mode, double scale, double bloom_threshold, int bloom_passes) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (s->nx < 2 || s->ny < 2) return fail(FS_E_INVALID_ARG, "render needs at least 2 x 2 cells");
  if (!std::isfinite(mode) || !std::isfinite(scale) || !std::isfinite(bloom_threshold))
    return fail(FS_E_INVALID_ARG, "render parameters must be finite");
  if ((rc = set_device(s))) return rc;
  const int bw = s->nx >> 1, bh = s->ny >> 1;
  const size_t nb = (size_t)bw * (size_t)bh;
  if (!s->out_img) {
    if ((rc = dalloc(&s->out_img, s->n)) || (rc = dalloc(&s->bloom[0], nb)) || (rc = dalloc(&s->bloom[1], nb))) {
      dfree(s->out_img);  // all three or none
      dfree(s->bloom[0]);
      dfree(s->bloom[1]);
      s->out_img = s->bloom[0] = s->bloom[1] = nullptr;
      return rc;
    }
    // a render with no bloom pass leaves the bloom as it was: zero before the first
    HIPCHK(hipMemsetAsync(s->bloom[0], 0, nb * sizeof(fsk::Half4), s->stream));
    s->bloom_cur = 0;
  }
  fsk::k_render<<<fsk::vt_grid(s->nx, s->ny), BLK, 0, s->stream>>>(
      s->u[s->cur], s->v[s->cur], s->dye[0] ? s->dye[s->dye_cur] : nullptr, s->out_img, s->nx, s->ny,
      (float)s->dx, (float)s->dy, (float)scale, (float)mode);
  LAUNCHCHK();
  if (bloom_passes > 0) {
    fsk::k_bloom_bright<<<grid2(bw, bh), BLK, 0, s->stream>>>(s->out_img, s->bloom[0], bw, bh, s->nx,
                                                             (float)bloom_threshold);
    LAUNCHCHK();
    s->bloom_cur = 0;
    for (int k = 0; k < bloom_passes; k++)
      for (int axis = 0; axis < 2; axis++) {
        fsk::k_bloom_blur<<<grid2(bw, bh), BLK, 0, s->stream>>>(s->bloom[s->bloom_cur], s->bloom[1 - s->bloom_cur],
                                                               bw, bh, axis);
        LAUNCHCHK();
        s->bloom_cur = 1 - s->bloom_cur;
      }
  }
  return FS_OK;
}

int fs_step_cfl(void* h, double cfl) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!(cfl > 0) || !std::isfinite(cfl)) return fail(FS_E_INVALID_ARG, "cfl must be > 0");
  if ((rc = set_device(s))) return rc;
  if (gated_eligible(s)) return step_gated(s, true, cfl);
  if ((rc = consume_records(s, true))) return rc;
  if ((rc = enqueue_head(s, true, cfl, 0))) return rc;
  return do_step(s);
}

int fs_sync(void* h) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  return sync_stream(s);
}

int fs_max_divergence(void* h, double* out) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!out) return fail(FS_E_INVALID_ARG, "out is NULL");
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  if ((rc = enqueue_divergence(s))) return rc;  // the reference also refreshes its div texture
  double d[fsk::DIAG_W];
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
    if ((rc = ensure_rpart(s, iterations))) return rc;
    fs_config saved = s->cfg;
    s->cfg.tolerance = 0;
    s->cfg.pressure_iterations = iterations;
    s->cfg.guess = FS_GUESS_ZERO;
    int cyc = 0;
    rc = solve_multigrid(s, &cyc);
    s->cfg = saved;
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(p.data(), s->p64[s->pidx], s->n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  } else {
    int pi = 0;
    if ((rc = solve_flat(s, iterations, &pi))) return rc;
    std::vector<float> pf(s->n);
    HIPCHK(hipMemcpyAsync(pf.data(), s->p32[pi], s->n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
    if ((rc = sync_stream(s))) return rc;
    for (size_t k = 0; k < s->n; k++) p[k] = pf[k];
  }
  HIPCHK(hipMemcpyAsync(dv.data(), s->div, s->n * sizeof(float), hipMemcpyDeviceToHost, s->stream));
  if ((rc = sync_stream(s))) return rc;
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
  if ((rc = ensure_rpart(s, iterations))) return rc;
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
  if ((rc = set_device(s)) || (rc = sync_stream(s))) return rc;
  destroy_graphs(s);
  s->cfg.tolerance = tol;
  s->pred_known = false;
  return FS_OK;
}

int fs_set_nu(void* h, double nu) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!(nu >= 0) || !std::isfinite(nu)) return fail(FS_E_INVALID_ARG, "nu must be >= 0");
  fs_solver_t* s = (fs_solver_t*)h;
  if ((nu > 0) != (s->cfg.nu > 0)) {  // captured bodies hold the diffusion on/off switch
    if ((rc = set_device(s)) || (rc = sync_stream(s))) return rc;
    destroy_graphs(s);
  }
  s->cfg.nu = nu;
  return FS_OK;
}

int fs_set_confinement(void* h, double eps) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!std::isfinite(eps)) return fail(FS_E_INVALID_ARG, "confinement must be finite");
  fs_solver_t* s = (fs_solver_t*)h;
  if (eps != s->cfg.confinement) {  // captured bodies hold the strength
    if ((rc = set_device(s)) || (rc = sync_stream(s))) return rc;
    destroy_graphs(s);
  }
  s->cfg.confinement = eps;
  return FS_OK;
}

int fs_set_guess(void* h, int guess) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  const bool learned = guess == FS_GUESS_LEARNED || guess == FS_GUESS_LEARNED_COARSE ||
                       guess == FS_GUESS_LEARNED_QUARTER;
  if (guess != FS_GUESS_ZERO && guess != FS_GUESS_PREVIOUS && guess != FS_GUESS_EXTRAPOLATED && !learned)
    return fail(FS_E_INVALID_ARG, "unknown guess mode %d", guess);
  if (guess == FS_GUESS_EXTRAPOLATED && s->cfg.pressure_solver != FS_P_MULTIGRID)
    return fail(FS_E_UNSUPPORTED, "the extrapolated guess needs the multigrid solver");
  if (guess == FS_GUESS_EXTRAPOLATED && !s->p_hist) {
    if ((rc = set_device(s)) || (rc = sync_stream(s))) return rc;
    if ((rc = dalloc(&s->p_hist, s->n))) return rc;
    s->have_hist = false;
  }
  if (learned && !s->guess_net.loaded)
    return fail(FS_E_INVALID_ARG, "the learned guess needs its weights: call fs_load_guess first");
  if (guess == FS_GUESS_LEARNED_COARSE && s->nx % 256 != 0)
    return fail(FS_E_UNSUPPORTED, "the half-resolution learned guess needs a grid size that is a multiple of 256, got %d",
                s->nx);
  if (guess == FS_GUESS_LEARNED_QUARTER && s->nx % 512 != 0)
    return fail(FS_E_UNSUPPORTED,
                "the quarter-resolution learned guess needs a grid size that is a multiple of 512, got %d", s->nx);
  if (guess != s->cfg.guess) {
    if ((rc = set_device(s)) || (rc = sync_stream(s))) return rc;
    destroy_graphs(s);
    s->pred_known = false;
  }
  s->cfg.guess = guess;
  return FS_OK;
}

// Weights blob: "FSGN", uint32 version 1, uint32 layer count 15, then per layer
// uint32 cin, cout, k and FP32 weights [cout][k*k][cin] followed by FP32 bias[cout].
int fs_load_guess(void* h, const void* blob, size_t bytes) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!blob) return fail(FS_E_INVALID_ARG, "blob is NULL");
  if (s->cfg.pressure_solver != FS_P_MULTIGRID)
    return fail(FS_E_UNSUPPORTED, "the learned guess needs the multigrid solver");
  if (s->nx != s->ny || s->nx % 128 != 0)
    return fail(FS_E_UNSUPPORTED, "the learned guess needs a square grid whose size is a multiple of 128, got %dx%d",
                s->nx, s->ny);
  const unsigned char* b = (const unsigned char*)blob;
  size_t off = 0;
  auto take = [&](void* dst, size_t nb) -> bool {
    if (off + nb > bytes) return false;
    std::memcpy(dst, b + off, nb);
    off += nb;
    return true;
  };
  char magic[4];
  uint32_t ver = 0, nl = 0;
  if (!take(magic, 4) || std::memcmp(magic, "FSGN", 4) != 0 || !take(&ver, 4) || ver != 1 || !take(&nl, 4) || nl != 15)
    return fail(FS_E_INVALID_ARG, "not a version-1 guess network blob");
  if ((rc = set_device(s))) return rc;
  if ((rc = sync_stream(s))) return rc;
  // a learned mode stays selected only if the new weights load: until then (and after
  // a failed load) the solver falls back to the previous pressure
  const int mode = s->cfg.guess;
  const bool learned_mode =
      mode == FS_GUESS_LEARNED || mode == FS_GUESS_LEARNED_COARSE || mode == FS_GUESS_LEARNED_QUARTER;
  if (learned_mode) s->cfg.guess = FS_GUESS_PREVIOUS;
  release_guess(s);
  auto& G = s->guess_net;
  auto bail = [&](int code) {
    release_guess(s);
    return code;
  };
  for (int l = 0; l < 15; l++) {
    const GuessLayer& L = kGuessLayers[l];
    uint32_t dims[3];
    if (!take(dims, sizeof dims)) return bail(fail(FS_E_INVALID_ARG, "blob truncated at layer %d", l));
    if ((int)dims[0] != L.cin || (int)dims[1] != L.cout || (int)dims[2] != L.k)
      return bail(fail(FS_E_INVALID_ARG, "layer %d is %ux%u k%u, expected %dx%d k%d", l, dims[0], dims[1], dims[2],
                       L.cin, L.cout, L.k));
    size_t nw = (size_t)L.cout * L.k * L.k * L.cin;
    std::vector<float> w(nw), bias(L.cout);
    if (!take(w.data(), nw * 4) || !take(bias.data(), (size_t)L.cout * 4))
      return bail(fail(FS_E_INVALID_ARG, "blob truncated at layer %d", l));
    for (float v : w)
      if (!std::isfinite(v)) return bail(fail(FS_E_INVALID_ARG, "non-finite weight in layer %d", l));
    if ((rc = dalloc(&G.b[l], (size_t)L.cout))) return bail(rc);
    HIPCHK(hipMemcpy(G.b[l], bias.data(), (size_t)L.cout * 4, hipMemcpyHostToDevice));
    if (l == 0 || l == 14) {
      float** dst = l == 0 ? &G.w0 : &G.wout;
      if ((rc = dalloc(dst, nw))) return bail(rc);
      HIPCHK(hipMemcpy(*dst, w.data(), nw * 4, hipMemcpyHostToDevice));
    } else {
      std::vector<uint16_t> wb(nw);
      for (size_t k = 0; k < nw; k++) {  // FP32 -> BF16, round to nearest even
        uint32_t u;
        std::memcpy(&u, &w[k], 4);
        u += 0x7fffu + ((u >> 16) & 1u);
        wb[k] = (uint16_t)(u >> 16);
      }
      if ((rc = dalloc(&G.w[l], nw))) return bail(rc);
      HIPCHK(hipMemcpy(G.w[l], wb.data(), nw * 2, hipMemcpyHostToDevice));
    }
  }
  if (off != bytes) return bail(fail(FS_E_INVALID_ARG, "%zu trailing bytes after the last layer", bytes - off));
  // enc0a's and dec0b's outputs never leave their kernels (k_guess_enc0, the output
  // epilogue), so only layers 1..12 have activation tensors
  for (int a = 1; a < 13; a++) {
    const GuessLayer& L = kGuessLayers[a];
    size_t m = (size_t)(s->nx >> L.level);
    if ((rc = dalloc(&G.act[a], m * m * L.cout))) return bail(rc);
  }
  G.nblk = 1024;
  if ((rc = dalloc(&G.scales, 2))) return bail(rc);
  if ((rc = dalloc(&G.ppart, (size_t)G.nblk))) return bail(rc);
  if (s->nx % 256 == 0) {  // the coarse variants' buffers (sized for half resolution)
    const size_t mm = (size_t)(s->nx / 2) * (s->nx / 2);
    if ((rc = dalloc(&G.rc, mm)) || (rc = dalloc(&G.pc, mm)) || (rc = dalloc(&G.dc, mm)) ||
        (rc = dalloc(&G.cpart, 2 * mm / 256)))
      return bail(rc);
  }
  G.loaded = true;
  s->cfg.guess = mode;
  return FS_OK;
}

int fs_set_forcing(void* h, const fs_forcing* f) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!f) {
    if (s->forcing_on) {  // captured bodies hold the forcing launches
      if ((rc = set_device(s)) || (rc = sync_stream(s))) return rc;
      destroy_graphs(s);
    }
    s->forcing_on = false;
    s->mode_list = false;
    return FS_OK;
  }
  if (!std::isfinite(f->kf) || !std::isfinite(f->width) || !std::isfinite(f->amp) || !(f->drag >= 0) ||
      !std::isfinite(f->drag))
    return fail(FS_E_INVALID_ARG, "forcing parameters must be finite, drag >= 0");
  if ((rc = set_device(s)) || (rc = sync_stream(s))) return rc;
  if (!s->ftab_col) {
    if ((rc = dalloc(&s->ftab_col, (size_t)fsk::MAX_MODES * s->nx)) ||
        (rc = dalloc(&s->ftab_row, (size_t)fsk::MAX_MODES * s->ny)))
      return rc;
  }
  if (!s->forcing_on || s->mode_list) destroy_graphs(s);
  s->forcing = *f;
  s->forcing_on = true;
  s->mode_list = false;
  return FS_OK;
}

int fs_set_forcing_modes(void* h, const double* kx, const double* ky, const double* phase, int count,
                         double amplitude, double drag, double phase_drift) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (count < 0) return fail(FS_E_INVALID_ARG, "count must be >= 0, got %d", count);
  if (count > 0 && (!kx || !ky || !phase)) return fail(FS_E_INVALID_ARG, "kx, ky or phase is NULL");
  if (count > fsk::MAX_MODES)
    return fail(FS_E_UNSUPPORTED, "at most %d forcing modes, got %d", fsk::MAX_MODES, count);
  for (int m = 0; m < count; m++)
    if (!std::isfinite(kx[m]) || !std::isfinite(ky[m]) || !std::isfinite(phase[m]))
      return fail(FS_E_INVALID_ARG, "mode %d is not finite", m);
  if (!std::isfinite(amplitude) || !(drag >= 0) || !std::isfinite(drag) || !std::isfinite(phase_drift))
    return fail(FS_E_INVALID_ARG, "amplitude and phase drift must be finite, drag >= 0");
  if ((rc = set_device(s)) || (rc = sync_stream(s))) return rc;
  destroy_graphs(s);  // captured bodies hold the forcing launches (or their absence)
  if (count == 0) {   // setForcing(null) or an empty list disables it (src/gpu/solver.js:372)
    s->forcing_on = false;
    s->mode_list = false;
    return FS_OK;
  }
  if (!s->ftab_col) {
    if ((rc = dalloc(&s->ftab_col, (size_t)fsk::MAX_MODES * s->nx)) ||
        (rc = dalloc(&s->ftab_row, (size_t)fsk::MAX_MODES * s->ny)))
      return rc;
  }
  s->mkx.assign(kx, kx + count);
  s->mky.assign(ky, ky + count);
  s->mph.assign(phase, phase + count);
  s->m_amp = amplitude;
  s->m_drag = drag;
  s->m_drift = phase_drift;
  s->mode_list = true;
  s->forcing_on = true;
  return FS_OK;
}

int fs_set_solid(void* h, const fs_disk* d) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (!d) {
    s->solid_on = false;
    return FS_OK;
  }
  if (!std::isfinite(d->x) || !std::isfinite(d->y) || !(d->r > 0) || !std::isfinite(d->r) ||
      !std::isfinite(d->vx) || !std::isfinite(d->vy) || !std::isfinite(d->omega))
    return fail(FS_E_INVALID_ARG, "disk centre, velocity must be finite and radius > 0");
  if (d->passes < 1 || d->passes > 16) return fail(FS_E_INVALID_ARG, "passes must be in 1..16, got %d", d->passes);
  if (!(s->cfg.solid_mask_width > 0)) return fail(FS_E_INVALID_ARG, "solid_mask_width must be > 0");
  if (!s->solid_on) {
    // no step has coupled yet: the impulse of "the last step" is zero
    if ((rc = set_device(s))) return rc;
    if ((rc = fill_f64(s, s->solid_imp, 3, 0.0))) return rc;
  }
  s->solid = *d;
  s->solid_on = true;
  return FS_OK;
}

int fs_solid_impulse(void* h, double* out) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!out) return fail(FS_E_INVALID_ARG, "out is NULL");
  fs_solver_t* s = (fs_solver_t*)h;
  if (!s->solid_on) return fail(FS_E_INVALID_ARG, "no disk is set");
  if ((rc = set_device(s))) return rc;
  HIPCHK(hipMemcpyAsync(s->host, s->solid_imp, sizeof(double) * 3, hipMemcpyDeviceToHost, s->stream));
  if ((rc = sync_stream(s))) return rc;
  out[0] = s->host[0];
  out[1] = s->host[1];
  out[2] = s->host[2];
  return FS_OK;
}

int fs_get_stats(void* h, fs_stats* out) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!out) return fail(FS_E_INVALID_ARG, "out is NULL");
  fs_solver_t* s = (fs_solver_t*)h;
  if (s->dt_pending > 0 || s->resid_row_pending >= 0 || !s->inflight.empty()) {
    if ((rc = set_device(s))) return rc;
    if ((rc = sync_stream(s))) return rc;
  }
  out->t = s->t;
  out->steps = s->steps;
  out->last_vcycles = s->last_vcycles;
  out->total_vcycles = s->total_vcycles;
  out->last_residual = s->last_residual;
  out->last_dt = s->last_dt;
  out->gated_steps = s->gated_steps;
  out->gate_misses = s->gate_misses;
  out->graph_captures = s->graph_captures;
  out->graph_launches = s->graph_launches;
  out->predicted_vcycles = s->pred_known ? s->last_pred : -1;
  return FS_OK;
}

int fs_diagnostics(void* h, fs_diag* out) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!out) return fail(FS_E_INVALID_ARG, "out is NULL");
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  double d[fsk::DIAG_W];
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
  out->enstrophy = 0.5 * d[8] / N;
  if (out->nonfinite) return fail(FS_E_NONFINITE, "%d non-finite velocity faces", out->nonfinite);
  return FS_OK;
}

}  // extern "C"

// Scratch of fs_diagnostics_noalloc: the per-workgroup partials of a fixed number of
// striding workgroups, then their fold. Each GPU the library runs on holds its own
// copy until the process ends; every call on that GPU shares it.
namespace {
constexpr int NA_GROUPS = 1024;
constexpr int NA_MAX_DEVICES = 64;
std::mutex g_na_mutex;
double* g_na_scratch[NA_MAX_DEVICES] = {};  // device address of `scratch` per device
}  // namespace
static __device__ __align__(256) double scratch[(NA_GROUPS + 1) * fsk::DIAG_W];

extern "C" {

int fs_diagnostics_noalloc(const float* u, const float* v, int nx, int ny, double lx, double ly, double* out,
                           void* stream) {
  if (!u || !v || !out) return fail(FS_E_INVALID_ARG, "u, v or out is NULL");
  if (nx < 2 || ny < 2) return fail(FS_E_INVALID_ARG, "grid must be >= 2x2, got %dx%d", nx, ny);
  if ((int64_t)nx * ny > (int64_t)1 << 28)
    return fail(FS_E_INVALID_ARG, "grid %dx%d exceeds the maximum of 2^28 cells", nx, ny);
  if (!(lx > 0) || !(ly > 0) || !std::isfinite(lx) || !std::isfinite(ly))
    return fail(FS_E_INVALID_ARG, "domain size must be positive and finite");
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NA_MAX_DEVICES) return fail(FS_E_DEVICE, "device %d beyond %d", dev, NA_MAX_DEVICES);
  double* base = nullptr;
  {
    // the array's address on this device, looked up once with the device current; the
    // array's host-side name is not a device address
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_scratch[dev]) {
      void* p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(FS_E_DEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
      g_na_scratch[dev] = (double*)p;
    }
    base = g_na_scratch[dev];
  }
  hipStream_t st = (hipStream_t)stream;
  double* part = base;
  double* raw = base + (size_t)NA_GROUPS * fsk::DIAG_W;
  const double dx = lx / nx, dy = ly / ny;
  fsk::k_diag_strided<<<NA_GROUPS, BLK, 0, st>>>(u, v, nx, ny, 1.0 / dx, 1.0 / dy, part);
  LAUNCHCHK();
  fsk::k_diag_final<<<1, 256, 0, st>>>(part, NA_GROUPS, raw);
  LAUNCHCHK();
  fsk::k_diag_finish<<<1, 64, 0, st>>>(raw, (double)nx * ny, dx * dy, out);
  LAUNCHCHK();
  return FS_OK;
}

int fs_run_kernel(void* h, int kernel, double dt) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if ((rc = set_device(s))) return rc;
  if (kernel == FS_K_ADVECT_SL || kernel == FS_K_ADVECT_MC || kernel == FS_K_DIFFUSE || kernel == FS_K_FORCING ||
      kernel == FS_K_ADVECT_DYE || kernel == FS_K_CONFINE) {
    if (!(dt >= 0) || !std::isfinite(dt)) return fail(FS_E_INVALID_ARG, "dt must be finite and >= 0");
    if ((rc = enqueue_head(s, false, dt, 0, false))) return rc;
  }
  switch (kernel) {
    case FS_K_ADVECT_SL: {
      int saved = s->cfg.advection;
      s->cfg.advection = FS_ADV_SEMI_LAGRANGIAN;
      rc = enqueue_advection(s);
      s->cfg.advection = saved;
      break;
    }
    case FS_K_ADVECT_MC:
      if (s->cfg.advection != FS_ADV_MACCORMACK)
        return fail(FS_E_INVALID_ARG, "MacCormack scratch exists only on a maccormack solver");
      rc = enqueue_advection(s);
      break;
    case FS_K_DIVERGENCE: rc = enqueue_divergence(s); break;
    case FS_K_DIFFUSE: rc = enqueue_diffusion(s); break;
    case FS_K_PROJECT: {
      int cyc = 0;
      rc = enqueue_projection(s, &cyc);
      s->last_vcycles = cyc;
      break;
    }
    case FS_K_SOLID:
      if (!s->solid_on) return fail(FS_E_INVALID_ARG, "no disk is set");
      rc = enqueue_solid(s);
      break;
    case FS_K_FORCING:
      if (!s->forcing_on) return fail(FS_E_INVALID_ARG, "no forcing is set");
      rc = enqueue_forcing(s);
      break;
    case FS_K_ADVECT_DYE:
      if (!s->dye[0]) return fail(FS_E_INVALID_ARG, "no dye: upload one or splat some first");
      rc = enqueue_dye(s);
      break;
    case FS_K_CONFINE:
      if (!(s->cfg.confinement > 0)) return fail(FS_E_INVALID_ARG, "confinement is off (strength <= 0)");
      rc = enqueue_confinement(s);
      break;
    case FS_K_GRADIENT:
      if (s->cfg.pressure_solver != FS_P_MULTIGRID)
        return fail(FS_E_INVALID_ARG, "the FP64 gradient needs the multigrid pressure");
      rc = enqueue_gradient64(s);
      break;
    default: return fail(FS_E_INVALID_ARG, "unknown kernel %d", kernel);
  }
  if (rc) return rc;
  s->have_maxabs = s->have_maxabs && kernel == FS_K_GRADIENT;
  return sync_stream(s);
}

int fs_tile_report(void* h, char* buf, size_t len) {
  int rc = check_solver(h);
  if (rc) return rc;
  if (!buf || len == 0) return fail(FS_E_INVALID_ARG, "buffer is NULL or empty");
  fs_solver_t* s = (fs_solver_t*)h;
  std::string out;
  char line[256];
  for (int l = 0; s->fused && l < s->coarse_start; l++) {
    const Level& L = s->lv[l];
    for (int dir = 0; dir < 2; dir++) {
      snprintf(line, sizeof line, "level %d n=%d %s: %s", l, L.n,
               dir == 0 ? (l == 0 ? "down0" : "down") : (l == 0 && up2_level0(s) ? "up (with level 1)" : "up"),
               kTileName[L.tile[dir]]);
      out += line;
      for (int k = 0; k < GRID_TILE_KINDS; k++)
        if (L.tune_us[dir][k] > 0) {
          snprintf(line, sizeof line, " | %s %.2f us", kTileName[k], L.tune_us[dir][k]);
          out += line;
        }
      out += "\n";
    }
  }
  snprintf(buf, len, "%s", out.c_str());
  return FS_OK;
}

int fs_profile_enable(void* h, unsigned mask) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (mask >> PK_COUNT) return fail(FS_E_INVALID_ARG, "unknown profile kinds in mask 0x%x", mask);
  if ((rc = set_device(s))) return rc;
  if ((rc = sync_stream(s))) return rc;
  destroy_graphs(s);  // a profiled pass launches one by one; later captures start afresh
  s->prof.mask = mask;
  for (int k = 0; k < PK_COUNT; k++) {
    s->prof.ms[k] = 0;
    s->prof.launches[k] = 0;
    s->prof.bytes[k] = 0;
  }
  return FS_OK;
}

int fs_profile_read(void* h, int kind, double* total_ms, int64_t* launches, double* bytes) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  if (kind < 0 || kind >= PK_COUNT) return fail(FS_E_INVALID_ARG, "unknown profile kind %d", kind);
  if ((rc = set_device(s))) return rc;
  if ((rc = sync_stream(s))) return rc;
  if (total_ms) *total_ms = s->prof.ms[kind];
  if (launches) *launches = s->prof.launches[kind];
  if (bytes) *bytes = s->prof.bytes[kind];
  return FS_OK;
}


int fs_energy_spectrum(void* h, double* E, int64_t* counts, int len) {
  int rc = check_solver(h);
  if (rc) return rc;
  fs_solver_t* s = (fs_solver_t*)h;
  const int n = s->nx;
  if (s->nx != s->ny || n < 4 || n > 4096 || (n & (n - 1)))
    return fail(FS_E_UNSUPPORTED, "the spectrum needs a square power-of-two grid up to 4096, got %dx%d", s->nx,
                s->ny);
  const int kmax = n / 2;
  if (len != kmax + 1) return fail(FS_E_INVALID_ARG, "E needs n/2 + 1 = %d entries, got %d", kmax + 1, len);
  if (!E) return fail(FS_E_INVALID_ARG, "E is NULL");
  if ((rc = set_device(s))) return rc;
  if (!s->spec_Z) {
    if ((rc = dalloc(&s->spec_Z, (size_t)n * n))) return rc;
    if ((rc = dalloc(&s->spec_part, (size_t)n * (kmax + 1)))) return rc;
    if ((rc = dalloc(&s->spec_E, (size_t)kmax + 1))) return rc;
  }
  int lg = 0;
  while ((1 << lg) < n) lg++;
  const int c = s->cur;
  size_t sh_rows = sizeof(fsk::cplx) * (size_t)(n + n / 2), sh_cols = sh_rows + sizeof(double) * (size_t)n;
  fsk::k_spec_rows<<<n, 256, sh_rows, s->stream>>>(s->u[c], s->v[c], s->spec_Z, n, lg);
  LAUNCHCHK();
  fsk::k_spec_cols<<<n, 256, sh_cols, s->stream>>>(s->spec_Z, s->spec_part, n, lg);
  LAUNCHCHK();
  fsk::k_spec_final<<<(unsigned)((kmax + 1 + 255) / 256), 256, 0, s->stream>>>(s->spec_part, s->spec_E, n);
  LAUNCHCHK();
  HIPCHK(hipMemcpyAsync(E, s->spec_E, sizeof(double) * (kmax + 1), hipMemcpyDeviceToHost, s->stream));
  if ((rc = sync_stream(s))) return rc;
  if (counts) {  // modes per shell (src/core/fft.js:214-226), independent of the field
    for (int k = 0; k <= kmax; k++) counts[k] = 0;
    for (int j = 0; j < n; j++) {
      long ky = j <= n / 2 ? j : j - n;
      for (int i = 0; i < n; i++) {
        long kx = i <= n / 2 ? i : i - n;
        int sh = (int)std::floor(std::sqrt((double)(kx * kx + ky * ky)) + 0.5);
        if (sh <= kmax) counts[sh]++;
      }
    }
  }
  return FS_OK;
}

// ------------------------------------------------------------ cavity batches

void fs_batch_config_default(fs_batch_config* c) {
  if (!c) return;
  c->advection = FS_ADV_MACCORMACK;
  c->tolerance = 1e-6;
  c->max_vcycles = 60;
  c->mg_pre = 2;
  c->mg_post = 2;
  c->coupling_passes = 2;
  c->mask_width = 1.5;
  c->guess = FS_GUESS_ZERO;
  c->device = 0;
  c->mg_fused = 1;
  c->graph = 1;
  c->mg_coarse_exact = 0;
  c->autotune = 1;
  c->tile = 0;
}

int fs_batch_create(const fs_batch_config* cfg, const fs_cavity* sims, int count, void** out) {
  if (!cfg || !sims || !out) return fail(FS_E_INVALID_ARG, "NULL argument");
  *out = nullptr;
  const fs_batch_config& c = *cfg;
  if (count < 1) return fail(FS_E_INVALID_ARG, "a batch needs at least one cavity");
  if (c.advection != FS_ADV_SEMI_LAGRANGIAN && c.advection != FS_ADV_MACCORMACK)
    return fail(FS_E_INVALID_ARG, "unknown advection scheme %d", c.advection);
  if (!(c.tolerance > 0)) return fail(FS_E_INVALID_ARG, "tolerance must be > 0");
  if (c.max_vcycles < 1 || c.max_vcycles > 1000) return fail(FS_E_INVALID_ARG, "max_vcycles must be in 1..1000");
  if (c.mg_pre < 0 || c.mg_pre > 8 || c.mg_post < 0 || c.mg_post > 8)
    return fail(FS_E_INVALID_ARG, "mg_pre and mg_post must be in 0..8");
  if (c.coupling_passes < 1 || c.coupling_passes > 16) return fail(FS_E_INVALID_ARG, "coupling_passes must be in 1..16");
  if (!(c.mask_width > 0)) return fail(FS_E_INVALID_ARG, "mask_width must be > 0");
  if (c.guess != FS_GUESS_ZERO && c.guess != FS_GUESS_PREVIOUS)
    return fail(FS_E_INVALID_ARG, "a batch takes the zero or the previous pressure guess");
  if (c.autotune < 0 || c.autotune > 2) return fail(FS_E_INVALID_ARG, "autotune must be 0, 1 or 2");
  if (c.tile < 0 || c.tile > GRID_TILE_KINDS) return fail(FS_E_INVALID_ARG, "tile must be in 0..%d", GRID_TILE_KINDS);
  if (c.mg_coarse_exact != 0 && c.mg_coarse_exact != 1)
    return fail(FS_E_INVALID_ARG, "mg_coarse_exact must be 0 or 1, got %d", c.mg_coarse_exact);
  for (int k = 0; k < count; k++) {
    const fs_cavity& q = sims[k];
    if (q.n < 64 || q.n > 4096 || (q.n & (q.n - 1)))
      return fail(FS_E_INVALID_ARG, "cavity %d: n must be a power of two in 64..4096, got %d", k, q.n);
    if (!std::isfinite(q.lid) || !(q.nu >= 0) || !std::isfinite(q.nu))
      return fail(FS_E_INVALID_ARG, "cavity %d: lid velocity must be finite and nu >= 0", k);
    if (q.has_disk && !(q.disk_r > 0 && std::isfinite(q.disk_x) && std::isfinite(q.disk_y)))
      return fail(FS_E_INVALID_ARG, "cavity %d: disk radius must be > 0 and its centre finite", k);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FS_E_DEVICE, "no HIP device");
  if (c.device < 0 || c.device >= ndev) return fail(FS_E_INVALID_ARG, "device %d of %d", c.device, ndev);
  fs_batch_t* b = new fs_batch_t();
  b->cfg = c;
  b->dev = c.device;
  b->fused = c.mg_fused && c.mg_pre >= 1 && c.mg_pre <= 3 && c.mg_post >= 1 && c.mg_post <= 3;
  auto bail = [&](int rc) {
    batch_release(b);
    return rc;
  };
  HIPCHK(hipSetDevice(b->dev));
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FS_E_DEVICE, "hipStreamCreate failed"));
  b->nsim = count;
  b->ext.resize(count);
  for (int k = 0; k < count; k++) b->ext[k] = k;
  std::stable_sort(b->ext.begin(), b->ext.end(), [&](int a, int bb) { return sims[a].n > sims[bb].n; });
  b->intl.resize(count);
  for (int k = 0; k < count; k++) b->intl[b->ext[k]] = k;
  int lg_max = 0;
  while ((1 << lg_max) < sims[b->ext[0]].n) lg_max++;
  b->max_lg = lg_max;
  // size classes, largest first
  long lvl_cells = 0;
  for (int lg = lg_max; lg >= 2; lg--) {
    long cnt = 0;
    for (int k = 0; k < count; k++)
      if (sims[b->ext[k]].n >= (1 << lg)) cnt++;
    b->count[lg] = cnt;
    b->base[lg] = lvl_cells;
    lvl_cells += cnt << (2 * lg);
  }
  b->sims_h.resize(count);
  std::vector<int> blk0(count);
  std::vector<long> cbase(count);
  long off = 0;
  for (int k = 0; k < count; k++) {
    const fs_cavity& q = sims[b->ext[k]];
    fsk::CavSim& S = b->sims_h[k];
    S.n = q.n;
    S.lg = 0;
    while ((1 << S.lg) < q.n) S.lg++;
    S.off = off;
    S.lvl = b->base[S.lg] + ((long)k << (2 * S.lg));
    S.lid = (float)q.lid;
    S.nu = (float)q.nu;
    S.disk = q.has_disk ? 1 : 0;
    S.dx_ = (float)q.disk_x;
    S.dy_ = (float)q.disk_y;
    S.dr = (float)q.disk_r;
    S.dt = 0;
    blk0[k] = (int)(off / 256);
    cbase[k] = b->base[6] + ((long)k << 12);
    off += (long)q.n * q.n;
  }
  b->cells = off;
  b->nblocks = (int)(off / 256);
  // after the first block of each cavity: the cavity of each block
  blk0.resize((size_t)count + b->nblocks);
  for (int k = 0; k < count; k++) {
    const int e = k + 1 < count ? blk0[k + 1] : b->nblocks;
    for (int q = blk0[k]; q < e; q++) blk0[(size_t)count + q] = k;
  }
  int rc = 0;
  for (int k = 0; k < 2 && !rc; k++) {
    rc = dalloc(&b->u[k], (size_t)off);
    if (!rc) rc = dalloc(&b->v[k], (size_t)off);
    if (!rc) rc = dalloc(&b->p64[k], (size_t)off);
  }
  if (!rc) rc = dalloc(&b->fu, (size_t)off);
  if (!rc) rc = dalloc(&b->fv, (size_t)off);
  if (!rc) rc = dalloc(&b->sims, (size_t)count);
  if (!rc) rc = dalloc(&b->blk0, blk0.size());
  if (!rc) rc = dalloc(&b->lx, (size_t)lvl_cells);
  if (!rc) rc = dalloc(&b->lr, (size_t)lvl_cells);
  if (!rc) rc = dalloc(&b->lxb, (size_t)lvl_cells);
  if (!rc) rc = dalloc(&b->cbase, (size_t)count);
  if (!rc) rc = dalloc(&b->pinv4, 256);
  if (!rc && c.mg_coarse_exact) rc = dalloc(&b->ex_tab, (size_t)fsk::EX_TAB * 4);
  b->rp_rows = c.max_vcycles + 2;
  if (!rc) rc = dalloc(&b->rpart, (size_t)b->rp_rows * b->nblocks);
  if (!rc) rc = halloc(&b->rpart_host, (size_t)b->rp_rows * b->nblocks);
  if (!rc) rc = dalloc(&b->mpart, (size_t)2 * b->nblocks);
  if (rc) return bail(rc);
  HIPCHK(hipMemcpy(b->sims, b->sims_h.data(), sizeof(fsk::CavSim) * count, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->blk0, blk0.data(), sizeof(int) * blk0.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->cbase, cbase.data(), sizeof(long) * count, hipMemcpyHostToDevice));
  std::vector<float> pinv = neumann_pinv4();
  HIPCHK(hipMemcpy(b->pinv4, pinv.data(), sizeof(float) * 256, hipMemcpyHostToDevice));
  if (b->ex_tab) {
    const std::vector<uint8_t> t = exact_coarse_table(64, 1.0 / 64, true);  // unit box, h = 1/64
    HIPCHK(hipMemcpy(b->ex_tab, t.data(), t.size(), hipMemcpyHostToDevice));
  }
  b->off = off;
  b->lvl_cells = lvl_cells;
  const int forced = c.tile >= 1 ? c.tile - 1 : -1;
  for (int lg = 0; lg < 13; lg++) b->tile[lg][0] = b->tile[lg][1] = forced >= 0 ? forced : TILE32;
  if ((rc = batch_reset(b))) return bail(rc);
  if (c.autotune >= 1 && forced < 0 && b->fused && (rc = batch_autotune_tiles(b))) return bail(rc);
  if (c.autotune >= 2 && b->fused && (rc = batch_autotune_sweeps(b))) return bail(rc);
  *out = b;
  return FS_OK;
}

int fs_batch_destroy(void* h) {
  batch_release((fs_batch_t*)h);
  return FS_OK;
}

int fs_batch_step_cfl(void* h, double cfl) {
  if (!h) return fail(FS_E_INVALID_ARG, "batch handle is NULL");
  fs_batch_t* b = (fs_batch_t*)h;
  if (!(cfl > 0) || !std::isfinite(cfl)) return fail(FS_E_INVALID_ARG, "cfl must be finite and > 0");
  HIPCHK(hipSetDevice(b->dev));
  int rc;
  const bool zero_p = !(b->cfg.guess == FS_GUESS_PREVIOUS && b->have_p);
  const int a0 = b->pidx;
  const int c = std::max(0, std::min(b->last_pred, b->cfg.max_vcycles));
  rc = b->cfg.graph ? batch_launch_head_graph(b, cfl, c, a0, zero_p) : batch_enqueue_head(b, cfl, c, a0, zero_p);
  if (rc) return rc;
  int cycles = 0;
  int prc = batch_finish_projection(b, c, a0, zero_p, &cycles);
  if (prc && prc != FS_E_NOT_CONVERGED) return prc;
  if ((rc = blaunch(b, fsk::k_cav_gradient64, dim3((b->nblocks + fsk::CAV_WPB - 1) / fsk::CAV_WPB),
                    dim3(64 * fsk::CAV_WPB), b->u[b->cur], b->v[b->cur], (const double*)b->p64[b->pidx],
                    (const fsk::CavSim*)b->sims, (const int*)b->blk0, b->nsim, b->mpart, b->nblocks)))
    return rc;
  b->have_p = true;
  b->last_vcycles = cycles;
  b->total_vcycles += cycles;
  b->steps++;
  return prc;
}

int fs_batch_sync(void* h) {
  if (!h) return fail(FS_E_INVALID_ARG, "batch handle is NULL");
  fs_batch_t* b = (fs_batch_t*)h;
  HIPCHK(hipSetDevice(b->dev));
  return bsync(b);
}

int fs_batch_read_velocity(void* h, int sim, float* u, float* v, size_t count) {
  if (!h || !u || !v) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_batch_t* b = (fs_batch_t*)h;
  if (sim < 0 || sim >= b->nsim) return fail(FS_E_INVALID_ARG, "cavity %d of %d", sim, b->nsim);
  const fsk::CavSim& S = b->sims_h[b->intl[sim]];
  if (count != (size_t)S.n * S.n) return fail(FS_E_INVALID_ARG, "cavity %d has %d cells, got %zu", sim, S.n * S.n, count);
  HIPCHK(hipSetDevice(b->dev));
  HIPCHK(hipMemcpyAsync(u, b->u[b->cur] + S.off, sizeof(float) * count, hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipMemcpyAsync(v, b->v[b->cur] + S.off, sizeof(float) * count, hipMemcpyDeviceToHost, b->stream));
  return bsync(b);
}

int fs_batch_read_dt(void* h, double* dt, int count) {
  if (!h || !dt) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_batch_t* b = (fs_batch_t*)h;
  if (count != b->nsim) return fail(FS_E_INVALID_ARG, "the batch has %d cavities, got %d", b->nsim, count);
  HIPCHK(hipSetDevice(b->dev));
  std::vector<fsk::CavSim> t(b->nsim);
  HIPCHK(hipMemcpyAsync(t.data(), b->sims, sizeof(fsk::CavSim) * b->nsim, hipMemcpyDeviceToHost, b->stream));
  int rc = bsync(b);
  if (rc) return rc;
  for (int k = 0; k < b->nsim; k++) dt[b->ext[k]] = t[k].dt;
  return FS_OK;
}

int fs_batch_get_stats(void* h, fs_batch_stats* out) {
  if (!h || !out) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_batch_t* b = (fs_batch_t*)h;
  out->steps = b->steps;
  out->cells = b->cells;
  out->last_vcycles = b->last_vcycles;
  out->total_vcycles = b->total_vcycles;
  out->last_residual = b->last_residual;
  out->graph_captures = b->graph_captures;
  out->graph_replays = b->graph_replays;
  return FS_OK;
}

int fs_batch_tune_report(void* h, char* buf, size_t len) {
  if (!h || !buf || len == 0) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_batch_t* b = (fs_batch_t*)h;
  std::string r;
  char line[256];
  for (int lg = b->max_lg; lg > 6; lg--)
    for (int dir = 0; dir < 2; dir++) {
      snprintf(line, sizeof line, "class %d^2 (%ld grids) %s: %s", 1 << lg, b->count[lg], dir ? "up" : "down",
               kTileName[b->tile[lg][dir]]);
      r += line;
      if (b->tune_us[lg][dir][0] > 0) {
        r += " [";
        for (int k = 0; k < GRID_TILE_KINDS; k++) {
          snprintf(line, sizeof line, "%s%s %.1f us", k ? ", " : "", kTileName[k], b->tune_us[lg][dir][k]);
          r += line;
        }
        r += "]";
      }
      r += "\n";
    }
  snprintf(line, sizeof line, "sweeps: pre %d, post %d\n", b->cfg.mg_pre, b->cfg.mg_post);
  r += line;
  for (const auto& t : b->sweep_trials) {
    snprintf(line, sizeof line, "  trial pre %d post %d: %.3f ms/step, %d V-cycles in 3 steps\n", t.pre, t.post,
             t.ms_per_step, t.vcycles);
    r += line;
  }
  snprintf(buf, len, "%s", r.c_str());
  return r.size() < len ? FS_OK : fail(FS_E_INVALID_ARG, "report truncated: needs %zu bytes", r.size() + 1);
}

// ------------------------------------------------------------ guess trainer

int fs_trainer_create(int device, int n, int batch, double h, void** out) {
  if (!out) return fail(FS_E_INVALID_ARG, "NULL argument");
  *out = nullptr;
  if (n < 128 || n > 4096 || n % 128 != 0)
    return fail(FS_E_UNSUPPORTED, "the trainer needs n a multiple of 128 in 128..4096, got %d", n);
  if (batch < 1 || batch > 256) return fail(FS_E_INVALID_ARG, "batch must be in 1..256, got %d", batch);
  if (!(h > 0) || !std::isfinite(h)) return fail(FS_E_INVALID_ARG, "h must be finite and > 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(FS_E_DEVICE, "no HIP device");
  if (device < 0 || device >= ndev) return fail(FS_E_INVALID_ARG, "device %d of %d", device, ndev);
  fs_trainer_t* t = new fs_trainer_t();
  t->dev = device;
  t->n = n;
  t->B = batch;
  t->h = h;
  auto bail = [&](int rc) {
    trainer_release(t);
    return rc;
  };
  HIPCHK(hipSetDevice(device));
  if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(FS_E_DEVICE, "hipStreamCreate failed"));
  size_t off = 0;
  for (int l = 0; l < 15; l++) {
    const GuessLayer& L = kGuessLayers[l];
    t->woff[l] = off;
    off += (size_t)L.cout * L.k * L.k * L.cin;
    t->boff[l] = off;
    off += (size_t)L.cout;
  }
  t->P = off;
  int rc;
  for (float** q : {&t->par, &t->grad, &t->m1, &t->m2})
    if ((rc = dalloc(q, t->P))) return bail(rc);
  for (int l = 1; l < 14; l++) {
    const GuessLayer& L = kGuessLayers[l];
    if ((rc = dalloc(&t->wf[l], (size_t)L.cout * 9 * L.cin)) || (rc = dalloc(&t->wt[l], (size_t)L.cout * 9 * L.cin)))
      return bail(rc);
  }
  const size_t nn = (size_t)n * n;
  for (int a = 0; a < 14; a++) {
    const GuessLayer& L = kGuessLayers[a];
    size_t m = (size_t)(n >> L.level);
    if ((rc = dalloc(&t->act[a], m * m * L.cout * batch)) || (rc = dalloc(&t->gz[a], m * m * L.cout * batch)))
      return bail(rc);
  }
  // skip gradients of enc0b (16 at n), enc1b (32 at n/2), enc2b (64 at n/4)
  if ((rc = dalloc(&t->gcat[0], nn * 16 * batch)) || (rc = dalloc(&t->gcat[1], nn / 4 * 32 * batch)) ||
      (rc = dalloc(&t->gcat[2], nn / 16 * 64 * batch)))
    return bail(rc);
  t->pblk = (int)std::min<size_t>(256, (nn + 255) / 256);
  t->oblk = (int)((nn * batch + 255) / 256);
  if ((rc = dalloc(&t->pp, nn * batch)) || (rc = dalloc(&t->r32, nn * batch)) || (rc = dalloc(&t->tgt, nn * batch)) ||
      (rc = dalloc(&t->err, nn * batch)) ||
      (rc = dalloc(&t->scales, (size_t)2 * batch)) || (rc = dalloc(&t->mr, (size_t)t->pblk * batch)) ||
      (rc = dalloc(&t->mp, (size_t)t->pblk * batch)) || (rc = dalloc(&t->opart, (size_t)t->oblk * 17)) ||
      (rc = dalloc(&t->lpart, (size_t)t->oblk)) || (rc = dalloc(&t->loss_d, 1)) || (rc = dalloc(&t->idx_d, batch)) ||
      (rc = halloc(&t->loss_h, 1)) || (rc = halloc(&t->idx_h, batch)))
    return bail(rc);
  size_t need = (size_t)304 * train_first_grid(t);
  train_all_weights(t, true, &need);
  t->part_cap = need;
  // group sums: rows/TR_RC + 1 groups of at most the longest row (73792: dec2a's
  // weights and bias), or of the output layer's 17 over its oblk rows
  const size_t need2 = need / fsk::TR_RC + 304 + 73792 + (size_t)((t->oblk + fsk::TR_RC - 1) / fsk::TR_RC) * 17;
  if ((rc = dalloc(&t->part, need)) || (rc = dalloc(&t->part2, need2))) return bail(rc);
  *out = t;
  return FS_OK;
}

int fs_trainer_destroy(void* h) {
  trainer_release((fs_trainer_t*)h);
  return FS_OK;
}

int fs_trainer_param_count(void* h, size_t* count) {
  if (!h || !count) return fail(FS_E_INVALID_ARG, "NULL argument");
  *count = ((fs_trainer_t*)h)->P;
  return FS_OK;
}

// Parameters from a weights blob (the fs_load_guess format); clears the Adam state.
int fs_trainer_load(void* h, const void* blob, size_t bytes) {
  if (!h || !blob) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  const unsigned char* b = (const unsigned char*)blob;
  size_t off = 0;
  auto take = [&](void* dst, size_t nb) -> bool {
    if (off + nb > bytes) return false;
    std::memcpy(dst, b + off, nb);
    off += nb;
    return true;
  };
  char magic[4];
  uint32_t ver = 0, nl = 0;
  if (!take(magic, 4) || std::memcmp(magic, "FSGN", 4) != 0 || !take(&ver, 4) || ver != 1 || !take(&nl, 4) || nl != 15)
    return fail(FS_E_INVALID_ARG, "not a version-1 guess network blob");
  std::vector<float> flat(t->P);
  for (int l = 0; l < 15; l++) {
    const GuessLayer& L = kGuessLayers[l];
    uint32_t dims[3];
    if (!take(dims, sizeof dims)) return fail(FS_E_INVALID_ARG, "blob truncated at layer %d", l);
    if ((int)dims[0] != L.cin || (int)dims[1] != L.cout || (int)dims[2] != L.k)
      return fail(FS_E_INVALID_ARG, "layer %d is %ux%u k%u, expected %dx%d k%d", l, dims[0], dims[1], dims[2], L.cin,
                  L.cout, L.k);
    size_t nw = t->boff[l] - t->woff[l];
    if (!take(flat.data() + t->woff[l], (nw + L.cout) * 4)) return fail(FS_E_INVALID_ARG, "blob truncated at layer %d", l);
  }
  if (off != bytes) return fail(FS_E_INVALID_ARG, "%zu trailing bytes after the last layer", bytes - off);
  for (float v : flat)
    if (!std::isfinite(v)) return fail(FS_E_INVALID_ARG, "non-finite parameter in the blob");
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipMemcpyAsync(t->par, flat.data(), t->P * 4, hipMemcpyHostToDevice, t->stream));
  HIPCHK(hipMemsetAsync(t->m1, 0, t->P * 4, t->stream));
  HIPCHK(hipMemsetAsync(t->m2, 0, t->P * 4, t->stream));
  int rc = train_pack(t);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(t->stream));
  t->adam_steps = 0;
  t->loaded = true;
  t->have_batch = false;
  return FS_OK;
}

// The current parameters as a weights blob; *bytes is set to the size it needs.
int fs_trainer_export(void* h, void* blob, size_t capacity, size_t* bytes) {
  if (!h || !bytes) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  const size_t need = 12 + 15 * 12 + t->P * 4;
  *bytes = need;
  if (!blob) return FS_OK;
  if (capacity < need) return fail(FS_E_INVALID_ARG, "the blob needs %zu bytes, got %zu", need, capacity);
  if (!t->loaded) return fail(FS_E_INVALID_ARG, "no parameters loaded");
  std::vector<float> flat(t->P);
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipMemcpyAsync(flat.data(), t->par, t->P * 4, hipMemcpyDeviceToHost, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  unsigned char* o = (unsigned char*)blob;
  uint32_t head[2] = {1, 15};
  std::memcpy(o, "FSGN", 4);
  std::memcpy(o + 4, head, 8);
  size_t off = 12;
  for (int l = 0; l < 15; l++) {
    const GuessLayer& L = kGuessLayers[l];
    uint32_t dims[3] = {(uint32_t)L.cin, (uint32_t)L.cout, (uint32_t)L.k};
    std::memcpy(o + off, dims, 12);
    off += 12;
    size_t nb = (t->boff[l] + L.cout - t->woff[l]) * 4;
    std::memcpy(o + off, flat.data() + t->woff[l], nb);
    off += nb;
  }
  return FS_OK;
}

// Room for `count` pressure snapshots (FP64 n x n each, zeroed); replaces any held.
int fs_trainer_reserve(void* h, size_t count) {
  if (!h) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  if (count < 2) return fail(FS_E_INVALID_ARG, "a pair needs at least two snapshots");
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipStreamSynchronize(t->stream));
  dfree(t->snap);
  t->snap = nullptr;
  t->nsnap = 0;
  const size_t cells = (size_t)t->n * t->n;
  int rc = dalloc(&t->snap, count * cells);
  if (rc) return rc;
  HIPCHK(hipMemsetAsync(t->snap, 0, count * cells * sizeof(double), t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  t->nsnap = count;
  t->have_batch = false;
  return FS_OK;
}

// Snapshot `slot` from host memory (n x n FP64, row-major).
int fs_trainer_store(void* h, size_t slot, const double* p, size_t cells) {
  if (!h || !p) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  if (slot >= t->nsnap) return fail(FS_E_INVALID_ARG, "slot %zu of %zu reserved", slot, t->nsnap);
  if (cells != (size_t)t->n * t->n) return fail(FS_E_INVALID_ARG, "a snapshot has %d^2 cells, got %zu", t->n, cells);
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipStreamSynchronize(t->stream));
  HIPCHK(hipMemcpyAsync(t->snap + slot * cells, p, cells * sizeof(double), hipMemcpyHostToDevice, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  return FS_OK;
}

// Snapshot `slot` from a solver's current pressure, device to device (the solver is
// n x n on the trainer's device; its queued work finishes first).
int fs_trainer_store_pressure(void* h, size_t slot, void* solver) {
  if (!h || !solver) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  fs_solver_t* s = (fs_solver_t*)solver;
  if (slot >= t->nsnap) return fail(FS_E_INVALID_ARG, "slot %zu of %zu reserved", slot, t->nsnap);
  if (s->nx != t->n || s->ny != t->n) return fail(FS_E_INVALID_ARG, "the solver is %dx%d, the trainer %dx%d", s->nx, s->ny, t->n, t->n);
  if (s->dev != t->dev) return fail(FS_E_INVALID_ARG, "the solver is on device %d, the trainer on %d", s->dev, t->dev);
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipStreamSynchronize(t->stream));
  const size_t cells = (size_t)t->n * t->n;
  HIPCHK(hipMemcpyAsync(t->snap + slot * cells, s->p64[s->pidx], cells * sizeof(double), hipMemcpyDeviceToDevice,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return FS_OK;
}

// Forward and backward pass over a batch of pairs: image b is
// (p_prev, p) = (snapshot idx[b], snapshot idx[b] + 1). Leaves the loss and the
// gradients on the device.
int fs_trainer_batch(void* h, const int* idx, int count) {
  if (!h || !idx) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  if (!t->loaded) return fail(FS_E_INVALID_ARG, "load parameters first (fs_trainer_load)");
  if (count != t->B) return fail(FS_E_INVALID_ARG, "the trainer takes batches of %d pairs, got %d", t->B, count);
  for (int b = 0; b < count; b++)
    if (idx[b] < 0 || (size_t)idx[b] + 1 >= t->nsnap)
      return fail(FS_E_INVALID_ARG, "pair %d starts at snapshot %d of %zu", b, idx[b], t->nsnap);
  HIPCHK(hipSetDevice(t->dev));
  // the previous batch's kernels may still read the index buffer
  HIPCHK(hipStreamSynchronize(t->stream));
  std::memcpy(t->idx_h, idx, sizeof(int) * count);
  HIPCHK(hipMemcpyAsync(t->idx_d, t->idx_h, sizeof(int) * count, hipMemcpyHostToDevice, t->stream));
  int rc = train_forward_backward(t);
  if (rc) return rc;
  t->have_batch = true;
  return FS_OK;
}

// The last batch's loss (waits for it).
int fs_trainer_loss(void* h, double* loss) {
  if (!h || !loss) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  if (!t->have_batch) return fail(FS_E_INVALID_ARG, "no batch has been run");
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipMemcpyAsync(t->loss_h, t->loss_d, sizeof(double), hipMemcpyDeviceToHost, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  *loss = *t->loss_h;
  return FS_OK;
}

// The last batch's gradient, flat in blob order (count = fs_trainer_param_count).
int fs_trainer_read_gradients(void* h, float* grad, size_t count) {
  if (!h || !grad) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  if (!t->have_batch) return fail(FS_E_INVALID_ARG, "no batch has been run");
  if (count != t->P) return fail(FS_E_INVALID_ARG, "the network has %zu parameters, got %zu", t->P, count);
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipMemcpyAsync(grad, t->grad, count * 4, hipMemcpyDeviceToHost, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  return FS_OK;
}

// The last batch's output minus target, [batch][n][n].
int fs_trainer_read_error(void* h, float* err, size_t count) {
  if (!h || !err) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  if (!t->have_batch) return fail(FS_E_INVALID_ARG, "no batch has been run");
  const size_t need = (size_t)t->n * t->n * t->B;
  if (count != need) return fail(FS_E_INVALID_ARG, "the batch has %zu pixels, got %zu", need, count);
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipMemcpyAsync(err, t->err, count * 4, hipMemcpyDeviceToHost, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  return FS_OK;
}

// The parameters, flat in blob order.
int fs_trainer_read_params(void* h, float* params, size_t count) {
  if (!h || !params) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  if (count != t->P) return fail(FS_E_INVALID_ARG, "the network has %zu parameters, got %zu", t->P, count);
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipMemcpyAsync(params, t->par, count * 4, hipMemcpyDeviceToHost, t->stream));
  HIPCHK(hipStreamSynchronize(t->stream));
  return FS_OK;
}

// One Adam update from the last batch's gradient, then the BF16 weight copies.
int fs_trainer_adam(void* h, double lr, double beta1, double beta2, double eps) {
  if (!h) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  if (!t->have_batch) return fail(FS_E_INVALID_ARG, "no batch has been run");
  if (!(lr >= 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
    return fail(FS_E_INVALID_ARG, "Adam needs lr >= 0, 0 <= beta < 1, eps > 0");
  HIPCHK(hipSetDevice(t->dev));
  t->adam_steps++;
  const double bc1 = 1.0 - std::pow(beta1, (double)t->adam_steps), bc2 = 1.0 - std::pow(beta2, (double)t->adam_steps);
  int rc = tlaunch(t, fsk::k_tr_adam, dim3((unsigned)((t->P + 255) / 256)), dim3(256), t->par, (const float*)t->grad,
                   t->m1, t->m2, t->P, (float)(lr / bc1), (float)(1.0 / std::sqrt(bc2)), (float)beta1, (float)beta2,
                   (float)eps);
  if (rc) return rc;
  return train_pack(t);
}

// Waits for the trainer's queued work.
int fs_trainer_sync(void* h) {
  if (!h) return fail(FS_E_INVALID_ARG, "NULL argument");
  fs_trainer_t* t = (fs_trainer_t*)h;
  HIPCHK(hipSetDevice(t->dev));
  HIPCHK(hipStreamSynchronize(t->stream));
  return FS_OK;
}

}  // extern "C"
