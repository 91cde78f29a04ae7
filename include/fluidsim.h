/*
 * fluidsim.h -- C ABI of the MI355X stable-fluids solver (libfluidsim.so).
 *
 * Plain C: no torch or HIP types cross this boundary. Every entry point returns
 * an int status (FS_OK = 0, negative on failure) and sets a thread-local message
 * readable with fs_last_error(). A handle is not thread-safe: callers serialise
 * calls per handle (the N-API shim queues them), handles on different threads
 * are independent.
 *
 * Each declaration names the reference interface it replaces (abho7/fluid-sim,
 * src/gpu/solver.js unless stated otherwise).
 */
#ifndef FLUIDSIM_H
#define FLUIDSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_ABI_VERSION 3

/* status codes */
enum {
  FS_OK = 0,
  FS_E_INVALID_ARG = -1,   /* bad size, selector or pointer */
  FS_E_DEVICE = -2,        /* a HIP call failed (message names it) */
  FS_E_OOM = -3,           /* device allocation failed */
  FS_E_NONFINITE = -4,     /* NaN/Inf found by an FP64 reduction */
  FS_E_NOT_CONVERGED = -5, /* tolerance not reached within the V-cycle cap */
  FS_E_UNSUPPORTED = -6    /* valid request this build does not implement */
};

/* advection selector: GPUFluidSolver({advection}) src/gpu/solver.js:64-72, :604-647 */
enum { FS_ADV_SEMI_LAGRANGIAN = 0, FS_ADV_MACCORMACK = 1 };

/* pressure selector: GPUFluidSolver({pressureSolver}) src/gpu/solver.js:686-738.
 * Unknown selectors are rejected (the reference silently falls back to Jacobi,
 * src/gpu/solver.js:725; the CPU solver throws, src/cpu/solver.js:288-290). */
enum { FS_P_JACOBI = 0, FS_P_RED_BLACK = 1, FS_P_MULTIGRID = 2 };

/* multigrid smoother storage precision (accumulation is always FP32) */
enum { FS_MG_FP32 = 0, FS_MG_FP16 = 1, FS_MG_BF16 = 2 };

/* initial pressure guess for the multigrid projection */
/* FS_GUESS_LEARNED_COARSE / FS_GUESS_LEARNED_QUARTER: the learned guess evaluated at
 * half / quarter resolution on the 2x2 / 4x4 averages of its inputs, its increment
 * interpolated bilinearly (grid size a multiple of 256 / 512); no counterpart in the
 * reference, which has no learned guess. */
/* FS_GUESS_EXTRAPOLATED: the last two pressures extrapolated linearly in time,
 * 2 p_n - p_{n-1} (the previous pressure until two exist); no counterpart in the
 * reference either. */
enum {
  FS_GUESS_ZERO = 0,
  FS_GUESS_PREVIOUS = 1,
  FS_GUESS_LEARNED = 2,
  FS_GUESS_LEARNED_COARSE = 3,
  FS_GUESS_LEARNED_QUARTER = 4,
  FS_GUESS_EXTRAPOLATED = 5
};

/* readable fields for fs_read_field */
enum {
  FS_F_U = 0,        /* float[nx*ny], current u (left faces) */
  FS_F_V = 1,        /* float[nx*ny], current v (bottom faces) */
  FS_F_DIV = 2,      /* float[nx*ny], divergence written by the last divergence pass */
  FS_F_PRESSURE = 3, /* double[nx*ny], pressure of the last projection */
  FS_F_RESIDUAL = 4, /* float[nx*ny], level-0 multigrid right-hand side of the last cycle */
  FS_F_DYE = 5,      /* float[nx*ny*4], rgba dye at the cell centres (after fs_upload_dye or a dye splat) */
  FS_F_OUTPUT = 6,   /* uint16[nx*ny*4], the last fs_render image: rgba IEEE half floats (the
                        reference's `output`, rgba16float; its readTexture refuses that format) */
  FS_F_BLOOM = 7     /* uint16[(nx/2)*(ny/2)*4], the last bloom pass's half-resolution rgba
                        half floats (the reference's bloom[0]) */
};

/* single kernels for per-kernel parity tests (tools/gputest.html:90-100, :121-129) */
enum {
  FS_K_ADVECT_SL = 0,   /* velocity <- semi-Lagrangian(velocity), dt */
  FS_K_ADVECT_MC = 1,   /* velocity <- limited MacCormack(velocity), dt */
  FS_K_DIVERGENCE = 2,  /* div <- divergence(velocity) */
  FS_K_DIFFUSE = 3,     /* velocity <- explicit diffusion(velocity), dt, nu */
  FS_K_PROJECT = 4,     /* one projection with the configured solver */
  FS_K_GRADIENT = 5,    /* velocity <- velocity - grad(pressure) */
  FS_K_SOLID = 6,       /* velocity <- direct forcing toward the disk's rigid motion */
  FS_K_FORCING = 7,     /* velocity <- forcing and drag of the current step index and time, dt */
  FS_K_ADVECT_DYE = 8,  /* dye <- semi-Lagrangian(dye) through the velocity, faded and clamped, dt */
  FS_K_CONFINE = 9      /* velocity <- velocity + vorticity-confinement force, dt, the configured strength */
};

/* GPUFluidSolver constructor options, src/gpu/solver.js:64-98 */
typedef struct fs_config {
  int nx, ny;                 /* cells; reference default n = 256 */
  double lx, ly;              /* domain size; default 2*pi */
  double nu;                  /* kinematic viscosity; default 0 */
  int advection;              /* FS_ADV_*; default MacCormack */
  int pressure_solver;        /* FS_P_*; default Jacobi */
  int pressure_iterations;    /* sweeps (Jacobi, red-black) or V-cycles; default 40.
                                 With tolerance > 0 it caps the V-cycles per step. */
  double tolerance;           /* > 0 (multigrid only): iterate until the FP64 residual
                                 max|div u* - L p| <= tolerance; 0 = fixed count */
  int mg_min_size;            /* stop coarsening below this size; default 8 */
  int mg_pre, mg_post;        /* red-black sweeps per level; default 2, 2 */
  int mg_coarse;              /* sweeps on the coarsest level; default 40 */
  int mg_precision;           /* FS_MG_*; default FP32 */
  int guess;                  /* FS_GUESS_*; default zero (the reference zeroes p) */
  double solid_mask_width;    /* in cells; default 1.5 */
  int device;                 /* HIP device ordinal; default 0 */
  int mg_fused;               /* 1: LDS-tiled V-cycle (two launches per level, levels <= 64
                                 in one workgroup); 0: one launch per operator. Same bits. */
  int autotune;               /* 1 (default): time each tile shape of the V-cycle kernels at
                                 creation and keep the fastest per level (same bits) */
  int adv_fused;              /* 1 (default): MacCormack advection and diffusion in one tiled
                                 launch; 0: three launches (forward, correction, diffusion).
                                 Same bits. */
  int gated;                  /* 1 (default): with a tolerance on the fused multigrid, the
                                 convergence check runs on the device and steps never wait for
                                 the host (see fs_stats); 0: every projection reads its residuals
                                 back and adds V-cycles until the gate is met. Same bits when
                                 both run the same V-cycles. */
  int graph;                  /* 1: a gated step's launches replay as a captured hipGraph (one
                                 per V-cycle count, buffer parity and initial-pressure kind,
                                 captured on first use or a step ahead); 0 (default): launched one
                                 by one, which measured 2.5 % faster per step at 1024^2 (the host
                                 stays ahead of the device either way). Same bits. */
  double dye_fade;            /* dyeFade: the dye is multiplied by 1 - dye_fade each step; default 0 */
  int mg_coarse_exact;        /* 0 (default): the reference's V-cycle down to mg_min_size with
                                 mg_coarse sweeps on the coarsest level; 1: the hierarchy stops at
                                 its first level of at most 64 x 64 cells, which is solved exactly
                                 (real Fourier basis, BF16 head + tail products on the matrix
                                 cores, ~1e-5 relative) - not in the reference; changes the
                                 bits, not the tolerance the gate enforces */
  int tile;                   /* 0 (default): the V-cycle tile shapes as the autotuner picks them;
                                 1..7: every grid-wide level uses 64x32, 32x32, 16x16, 8x8, 64x64
                                 tiles, or 32x32 with 1024 / 16x16 with 256 threads (no tuning).
                                 Same bits. */
  double confinement;         /* confinement: vorticity-confinement strength (src/gpu/solver.js:67,
                                 :558-568); the force runs in each step when > 0; default 0 */
} fs_config;

/* one splat of step(dt, {splats}) (src/gpu/solver.js:546-554, _writeSplat :745-757): a
 * Gaussian of `radius` about (x, y) adding (dx, dy) g to the velocity, or with is_dye
 * (r, g, b) g to the dye */
typedef struct fs_splat {
  double x, y;                /* centre, world units */
  double dx, dy;              /* velocity impulse (force splats) */
  double r, g, b;             /* colour (dye splats) */
  double radius;              /* Gaussian radius; the reference's default is 0.2 */
  int is_dye;
} fs_splat;

/* per-step statistics (new: the reference reports none) */
typedef struct fs_stats {
  double t;                   /* simulated time */
  int64_t steps;              /* steps taken */
  int last_vcycles;           /* V-cycles in the last projection */
  int64_t total_vcycles;      /* V-cycles since creation */
  double last_residual;       /* FP64 max|div u* - L p| at the end of the last projection */
  double last_dt;             /* dt of the last step */
  /* Device-gated projection: with a tolerance on the fused multigrid, a step enqueues
   * the V-cycle count of the last completed step plus one spare, each V-cycle skipped
   * on the device once the FP64 residual gate is met, so no step waits for the host.
   * A step that needed more than it enqueued ends above the tolerance and is counted in
   * gate_misses (the next steps enqueue more); the first step after an upload or a
   * change of guess or tolerance is checked on the host and sets the prediction. */
  int64_t gated_steps;        /* steps projected that way */
  int64_t gate_misses;        /* of them, steps whose final residual is above the tolerance */
  int64_t graph_captures;     /* step bodies captured as hipGraphs */
  int64_t graph_launches;     /* steps whose body was launched as a graph */
  int predicted_vcycles;      /* V-cycles the next gated step expects (-1: none yet) */
} fs_stats;

/* FP64 diagnostics of the current velocity (src/core/grid.js:290-319) */
typedef struct fs_diag {
  double kinetic_energy;      /* 1/2 <u_c^2 + v_c^2>, centre-interpolated (grid.js:290-295) */
  double max_divergence;      /* max |div u| of the stored FP32 field, in FP64 */
  double l2_divergence;       /* sqrt(<div^2>) */
  double momentum_x, momentum_y; /* sum u dA, sum v dA (src/cpu/solid.js:264-270) */
  double max_abs_u, max_abs_v;
  int nonfinite;              /* count of non-finite faces */
  double enstrophy;           /* 1/2 <omega_c^2>, omega_c the centre average of the corner
                                 vorticities (grid.js:228-257, :297-303) */
} fs_diag;

/* band forcing (src/cpu/solver.js:179-237 semantics: force, then drag) */
typedef struct fs_forcing {
  double kf, width, amp;      /* shell centre, half width, amplitude */
  double drag;                /* linear drag alpha; u *= exp(-alpha dt) after the force */
  uint32_t seed;              /* seed of step k is seed + k */
} fs_forcing;

/* immersed rigid disk for one step (RigidDisk, src/cpu/solid.js:52-66): centre,
 * radius, rigid velocity; `passes` direct-forcing passes (the CPU default is 2,
 * src/cpu/solver.js:258; the reference GPU does 1). The mask width is
 * fs_config.solid_mask_width cells. */
typedef struct fs_disk {
  double x, y, r;
  double vx, vy, omega;
  int passes;
} fs_disk;

/* one lid-driven cavity of a batch (new: the reference's domain is periodic and has
 * no walls). Unit box, n x n cells (a power of two, 64..4096), no slip on every wall,
 * the top wall moving with velocity `lid`, viscosity nu, optionally a pinned disk. */
typedef struct fs_cavity {
  int n;
  double lid, nu;
  int has_disk;
  double disk_x, disk_y, disk_r;
} fs_cavity;

typedef struct fs_batch_config {
  int advection;              /* FS_ADV_*; default MacCormack */
  double tolerance;           /* FP64 residual gate for every cavity; default 1e-6 */
  int max_vcycles;            /* cap per step; default 60 */
  int mg_pre, mg_post;        /* red-black sweeps; default 2, 2 */
  int coupling_passes;        /* disk forcing passes; default 2 (the CPU default) */
  double mask_width;          /* disk mask width in cells; default 1.5 */
  int guess;                  /* FS_GUESS_ZERO (default) or FS_GUESS_PREVIOUS */
  int device;
  int mg_fused;               /* 1 (default): LDS-tiled level kernels; 0: one launch per operator.
                                 Same bits. */
  int graph;                  /* 1 (default): each step's launches up to the projection's read-back
                                 replay as a captured hipGraph (one per predicted V-cycle count,
                                 pressure parity and cfl); 0: launched one by one. Same bits. */
  int autotune;               /* 0: 32x32 tiles, the configured sweeps; 1 (default): the tile shape of
                                 each size class's level kernels timed at create (same bits);
                                 2: also (pre, post) in {1,2,3}^2 timed over steps from rest at
                                 create (the choice changes the bits) */
  int tile;                   /* 0 (default): as autotune says; 1..7: every size class uses
                                 64x32, 32x32, 16x16, 8x8, 64x64 tiles, or 32x32 with 1024 /
                                 16x16 with 256 threads (no tile timing). Same bits. */
  int mg_coarse_exact;        /* 0 (default): every cavity's levels 64 .. 4 in one workgroup with
                                 a 4 x 4 pseudo-inverse; 1: its 64 x 64 level solved exactly in the
                                 cell-centred cosine (DCT-II) basis of the zero-flux Laplacian on
                                 the matrix cores, one workgroup per cavity */
} fs_batch_config;

typedef struct fs_batch_stats {
  int64_t steps;
  int64_t cells;              /* cells over all cavities */
  int last_vcycles;           /* V-cycles of the last projection (every cavity runs them) */
  int64_t total_vcycles;
  double last_residual;       /* max over cavities of the FP64 residual */
  int64_t graph_captures;     /* step graphs captured so far */
  int64_t graph_replays;      /* steps whose head ran as a graph */
} fs_batch_stats;

int         fs_abi_version(void);
const char* fs_last_error(void);

/* createDevice() src/gpu/solver.js:43-62: device count and name */
int fs_device_count(int* count);
int fs_device_name(int device, char* buf, size_t len);

/* reference defaults, src/gpu/solver.js:64-98 and src/gpu/multigrid.js:159-161 */
void fs_config_default(fs_config* cfg);

/* new GPUFluidSolver(device, opts) src/gpu/solver.js:64-98 */
int fs_solver_create(const fs_config* cfg, void** solver);
/* destroy() src/gpu/solver.js:898-911; NULL is accepted */
int fs_solver_destroy(void* solver);

/* uploadVelocity(u, v) src/gpu/solver.js:301-315; n = nx*ny values each */
int fs_upload_velocity_f64(void* solver, const double* u, const double* v, size_t n);
int fs_upload_velocity_f32(void* solver, const float* u, const float* v, size_t n);
/* readVelocity() src/gpu/solver.js:356-362 */
int fs_read_velocity(void* solver, float* u, float* v, size_t n);
/* readTexture(tex) src/gpu/solver.js:336-353, by field id; bytes must match the field */
int fs_read_field(void* solver, int field, void* out, size_t bytes);

/* step(dt) src/gpu/solver.js:518-591: enqueue one step on the solver's stream */
int fs_step(void* solver, double dt);
/* step(dt, {splats}) src/gpu/solver.js:518-591: the splats applied in order after the
 * forcing (SPLAT src/gpu/shaders.js:474-511), each with its own parameters (the reference
 * binds the first splat's for every pass, src/gpu/solver.js:546-553 against :757-768);
 * a dye splat allocates the dye. Once a dye exists every step ends by advecting it
 * (ADVECT_SCALAR src/gpu/shaders.js:234-266). */
int fs_step_splats(void* solver, double dt, const fs_splat* splats, int count);
/* uploadDye(rgba) src/gpu/solver.js:317-323: n = nx*ny*4 floats, rgba per cell centre */
int fs_upload_dye(void* solver, const float* rgba, size_t n);
/* render({mode, scale, bloomThreshold, bloomPasses}) src/gpu/solver.js:771-823 (defaults
 * 0, 1.5, 0.75, 2): the display image of the current velocity and dye (RENDER
 * src/gpu/shaders.js:677-780; mode < 0.5 colours the vorticity, else the speed), then,
 * when bloom_passes > 0, the bright pass at half resolution and bloom_passes rounds of
 * the separable blur in x and y (BLOOM_BRIGHT :783-812, BLOOM_BLUR :816-843). Enqueued
 * on the solver's stream; read back with FS_F_OUTPUT and FS_F_BLOOM. Needs nx, ny >= 2. */
int fs_render(void* solver, double mode, double scale, double bloom_threshold, int bloom_passes);
/* step with dt = cfl / (max|u|/dx + max|v|/dy), the CFL definition of
 * src/core/grid.js:322-327, evaluated on the device from the last projection */
int fs_step_cfl(void* solver, double cfl);
/* device.queue.onSubmittedWorkDone() */
int fs_sync(void* solver);

/* maxDivergence() src/gpu/solver.js:835-850 (FP64 reduction on the device) */
int fs_max_divergence(void* solver, double* out);
/* poissonResidualAfter(it, dt) src/gpu/solver.js:857-896: ||L p - div|| / ||div|| */
int fs_poisson_residual_after(void* solver, int iterations, double* out);

/* mutable fields (index.html:472 mutates pressureIterations) */
int fs_set_pressure_iterations(void* solver, int iterations);
int fs_set_tolerance(void* solver, double tolerance);
int fs_set_nu(void* solver, double nu);
/* the confinement field (index.html:471 mutates it); <= 0 turns the force off */
int fs_set_confinement(void* solver, double eps);
int fs_set_guess(void* solver, int guess);

/* learned initial pressure guess (new: the reference zeroes p every step,
 * src/gpu/solver.js:669-672). Loads the weights of the 4-level U-Net that maps the
 * residual of the previous pressure and the previous pressure to an increment;
 * blob: "FSGN", uint32 1, uint32 15, then per layer uint32 cin, cout, k, FP32
 * weights [cout][k*k][cin], FP32 bias[cout]. Needs the multigrid solver (FP32 or
 * BF16 storage) and a square grid whose size is a multiple of 128. Then
 * fs_set_guess(solver, FS_GUESS_LEARNED), or FS_GUESS_LEARNED_COARSE /
 * FS_GUESS_LEARNED_QUARTER for a size that is a multiple of 256 / 512. A failed load
 * releases the previous weights and leaves the solver on FS_GUESS_PREVIOUS. */
int fs_load_guess(void* solver, const void* blob, size_t bytes);

/* seeded band forcing with the CPU's semantics (bandForcing + linearDrag,
 * src/cpu/solver.js:179-237: 24 modes from seed + step, force then drag); NULL disables */
int fs_set_forcing(void* solver, const fs_forcing* forcing);
/* setForcing(modes, {amplitude, drag, phaseDrift}) src/gpu/solver.js:370-414 with the GPU
 * kernel's semantics (FORCING src/gpu/shaders.js:532-566): u <- u exp(-drag dt) + amplitude
 * dt / sqrt(count) * sum (ky/|k|, -kx/|k|) cos(kx x + ky y + phase + phaseDrift t), t the
 * simulated time at the start of the step. count = 0 disables (setForcing(null)); at most
 * 32 modes. Replaces a band forcing and is replaced by one. */
int fs_set_forcing_modes(void* solver, const double* kx, const double* ky, const double* phase, int count,
                         double amplitude, double drag, double phase_drift);

/* setSolid(disk) src/gpu/solver.js:425-440 with the state _encodeSolid uploads each
 * step (:442-452): the next steps couple this disk; NULL removes it. Callers that
 * move the disk call this again before each step. */
int fs_set_solid(void* solver, const fs_disk* disk);
/* the sum integrateSolid reads back (src/gpu/solver.js:454-510), reduced on the
 * device in FP64 for the most recent coupling: out[0..1] = fluid impulse
 * sum(du dA), sum(dv dA); out[2] = torque impulse about the centre */
int fs_solid_impulse(void* solver, double* out);

int fs_get_stats(void* solver, fs_stats* out);
int fs_diagnostics(void* solver, fs_diag* out);

/* energySpectrum(cu, cv) of src/core/fft.js:196-234 on the device, in FP64: the
 * centre-interpolated velocity, shells round(|k|) <= n/2, E[k] = sum of
 * 0.5 (|U_k|^2 + |V_k|^2) / N^2 over the shell (Parseval: sum E = kinetic energy less
 * the corner modes). Square power-of-two grids up to 4096; len = n/2 + 1; counts
 * (modes per shell) may be NULL. */
int fs_energy_spectrum(void* solver, double* E, int64_t* counts, int len);

/* No counterpart in the reference: a variant of fs_diagnostics for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). It reads caller-owned device fields u, v (nx*ny FP32 each, MAC order, on
 * the current HIP device) and writes 9 doubles to the caller-owned device buffer `out`:
 * {kinetic energy, max|div|, l2 div, momentum x, momentum y, max|u|, max|v|, non-finite
 * faces, enstrophy}, the definitions of fs_diag. Enqueued on `stream` (a hipStream_t;
 * NULL = the default stream), nothing is waited for. It allocates nothing: its partials
 * live in a fixed 74 KB device array per GPU, held until the process ends and shared by
 * every call on that GPU, so calls that overlap on one GPU race. At most 2^28 cells. */
int fs_diagnostics_noalloc(const float* u, const float* v, int nx, int ny, double lx, double ly, double* out,
                           void* stream);

/* per-kernel hook replacing gputest's use of internals */
int fs_run_kernel(void* solver, int kernel, double dt);

/* measurement (new: the reference allocates a timestamp query set and never writes
 * it, src/gpu/solver.js:202-210). Kernel classes: */
enum {
  FS_PK_ADVECT_SL = 0, FS_PK_ADVECT_MC = 1, FS_PK_DIFFUSE = 2, FS_PK_DOWN0 = 3, FS_PK_DOWN = 4,
  FS_PK_COARSE = 5, FS_PK_UP = 6, FS_PK_GRADIENT = 7, FS_PK_RESIDUAL = 8, FS_PK_RB = 9,
  FS_PK_GUESS = 10,  /* 3x3 layers of the learned guess */
  FS_PK_ADVECT_FUSED = 11,  /* MacCormack + diffusion in one launch */
  FS_PK_FORCING = 12,       /* band forcing + drag */
  FS_PK_UP_COARSE = 13,     /* two levels up with the exact 64^2 coarse solve inside the launch */
  FS_PK_COUNT = 14
};
/* Time every launch of the classes in `mask` (bit k = class k; 0 = off) with HIP
 * events recorded at the dispatch's own start and end on the solver's stream, and
 * accumulate each launch's algorithmic bytes (inputs read once, outputs written once). */
int fs_profile_enable(void* solver, unsigned mask);
/* the V-cycle tile shape chosen for each grid-wide level and the autotuner's timings, as text */
int fs_tile_report(void* solver, char* buf, size_t len);

/* batched cavities: every stage of a step is one launch over all of them; each
 * cavity advances with its own dt = cfl h / (max(max|u|, |lid|) + max|v|). Cavity
 * indices are the caller's order. */
void fs_batch_config_default(fs_batch_config* cfg);
int fs_batch_create(const fs_batch_config* cfg, const fs_cavity* cavities, int count, void** batch);
int fs_batch_destroy(void* batch);
int fs_batch_step_cfl(void* batch, double cfl);
int fs_batch_sync(void* batch);
int fs_batch_read_velocity(void* batch, int cavity, float* u, float* v, size_t count);
int fs_batch_read_dt(void* batch, double* dt, int count);   /* dt of the last step, per cavity */
int fs_batch_get_stats(void* batch, fs_batch_stats* out);
/* the tile shape per size class and the sweeps chosen, with the autotuner's timings, as text */
int fs_batch_tune_report(void* batch, char* buf, size_t len);
int fs_profile_read(void* solver, int kind, double* total_ms, int64_t* launches, double* bytes);

/* trainer of the learned guess (new: the reference has no learned component). The
 * network's forward pass is the guess's own kernels over a batch; the backward pass
 * and Adam are the library's kernels. Parameters are FP32 in blob order (per layer
 * w[cout][k*k][cin], b[cout]); the loss is the mean over the batch and the pixels
 * of (o - t)^2 with t = (p - p_prev) / (max|r| h^2), r = L p - L p_prev. n is a
 * multiple of 128, h the grid spacing of the pressure snapshots. */
int fs_trainer_create(int device, int n, int batch, double h, void** trainer);
int fs_trainer_destroy(void* trainer);
int fs_trainer_param_count(void* trainer, size_t* count);
/* parameters from a weights blob (fs_load_guess's format); clears the Adam moments */
int fs_trainer_load(void* trainer, const void* blob, size_t bytes);
/* the parameters as a weights blob; *bytes = the size needed (blob may be NULL) */
int fs_trainer_export(void* trainer, void* blob, size_t capacity, size_t* bytes);
/* the trainer's store of pressure snapshots (FP64 n x n each, device memory):
 * room for `count` (zeroed), one from host memory, one from a solver's current
 * pressure (device to device; same n and device) */
int fs_trainer_reserve(void* trainer, size_t count);
int fs_trainer_store(void* trainer, size_t slot, const double* p, size_t cells);
int fs_trainer_store_pressure(void* trainer, size_t slot, void* solver);
/* forward and backward over `count` (= batch) pairs (snapshot idx[b], snapshot idx[b] + 1) */
int fs_trainer_batch(void* trainer, const int* idx, int count);
int fs_trainer_loss(void* trainer, double* loss);                      /* of the last batch */
int fs_trainer_read_gradients(void* trainer, float* grad, size_t count);
int fs_trainer_read_params(void* trainer, float* params, size_t count);
int fs_trainer_read_error(void* trainer, float* err, size_t count);   /* o - t, [batch][n][n] */
/* one Adam step (PyTorch's formulation) from the last batch's gradient */
int fs_trainer_adam(void* trainer, double lr, double beta1, double beta2, double eps);
int fs_trainer_sync(void* trainer);

#ifdef __cplusplus
}
#endif
#endif
