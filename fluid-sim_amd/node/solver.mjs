// solver.mjs -- drop-in for the reference's src/gpu/solver.js, backed by the
// MI355X library through the Node-API addon (fluidsim.node -> libfluidsim.so).
//
// Exports the same two names (createDevice, GPUFluidSolver) with the same
// constructor options, defaults, methods and Promise-returning readbacks.
// Deliberate differences: unknown selectors throw (the reference silently runs
// Jacobi / semi-Lagrangian, src/gpu/solver.js:606, :725); splats, vorticity
// confinement, dye and rendering are display features that are not built;
// forcing is the seeded band forcing of src/cpu/solver.js:179-237 via
// setBandForcing(), and an explicit mode list passed to setForcing() throws.
// Written in Node 12 syntax (no optional chaining, no nullish coalescing).
import { createRequire } from "module";

const require = createRequire(import.meta.url);
const addon = require("./fluidsim.node");

const ADVECTION = { "semi-lagrangian": 0, maccormack: 1 };
const PRESSURE = { jacobi: 0, "red-black": 1, multigrid: 2 };
const GUESS = { zero: 0, previous: 1, learned: 2, "learned-coarse": 3, "learned-quarter": 4 };

export async function createDevice({ requireTimestamps = false } = {}) {
  const d = addon.device();
  if (!d.count) throw new Error("no HIP device available");
  const solvers = new Set();
  const device = {
    _solvers: solvers,
    queue: {
      // the reference awaits this after step() to time the GPU
      onSubmittedWorkDone: async () => {
        for (const s of solvers) s.sync();
      },
    },
  };
  return {
    device,
    hasTimestamps: true,
    info: { vendor: "amd", architecture: "gfx950", device: "", description: d.description },
  };
}

export class GPUFluidSolver {
  constructor(device, {
    n = 256, nx = n, ny = n, lx = 2 * Math.PI, ly = 2 * Math.PI,
    nu = 0.0, confinement = 0, dyeFade = 0,
    advection = "maccormack",
    solidMaskWidth = 1.5,
    pressureSolver = "jacobi", pressureIterations = 40,
    hasTimestamps = false,
    tolerance = 0, mgMinSize = 8, mgPre = 2, mgPost = 2, mgCoarse = 40, guess = "zero",
  } = {}) {
    if (!(advection in ADVECTION)) {
      throw new Error(`unknown advection scheme ${advection}; have ${Object.keys(ADVECTION).join(", ")}`);
    }
    if (!(pressureSolver in PRESSURE)) {
      throw new Error(`unknown pressure solver ${pressureSolver}; have ${Object.keys(PRESSURE).join(", ")}`);
    }
    if (!(guess in GUESS)) throw new Error(`unknown guess mode ${guess}`);
    if (confinement > 0) throw new Error("vorticity confinement is display-only and not built here");
    this.device = device;
    this.nx = nx; this.ny = ny;
    this.dx = lx / nx; this.dy = ly / ny;
    this._nu = nu;
    this.advection = advection;
    this.pressureSolver = pressureSolver;
    this._pressureIterations = pressureIterations;
    this.solidMaskWidth = solidMaskWidth;
    this.solid = null;
    this.dyeFade = dyeFade;
    this.hasTimestamps = true;
    this._h = addon.create({
      nx, ny, lx, ly, nu,
      advection: ADVECTION[advection],
      pressureSolver: PRESSURE[pressureSolver],
      pressureIterations, tolerance, mgMinSize, mgPre, mgPost, mgCoarse,
      guess: GUESS[guess], solidMaskWidth,
    });
    if (device && device._solvers) device._solvers.add(this);
  }

  get nu() { return this._nu; }
  set nu(v) { addon.setParam(this._h, "nu", v); this._nu = v; }

  // mutable after construction, as index.html:472 uses it
  get pressureIterations() { return this._pressureIterations; }
  set pressureIterations(v) { addon.setParam(this._h, "pressureIterations", v); this._pressureIterations = v; }

  get t() { return addon.stats(this._h).t; }
  get steps() { return addon.stats(this._h).steps; }

  setTolerance(tol) { addon.setParam(this._h, "tolerance", tol); }
  // the learned guess (no counterpart in the reference): loadGuess(weights blob), then
  // setGuess("learned" | "learned-coarse" | "learned-quarter")
  loadGuess(bytes) { addon.loadGuess(this._h, bytes); }
  setGuess(g) {
    if (!(g in GUESS)) throw new Error(`unknown guess mode ${g}`);
    addon.setParam(this._h, "guess", GUESS[g]);
  }

  uploadVelocity(u, v) { addon.uploadVelocity(this._h, u, v); }

  async readVelocity() { return addon.readVelocity(this._h, this.nx * this.ny); }

  step(dt, { splats = [] } = {}) {
    if (splats.length) throw new Error("splats are display-only and not built here");
    this._pushSolid();
    addon.step(this._h, dt);
  }

  stepCFL(cfl = 0.5) { this._pushSolid(); addon.stepCFL(this._h, cfl); }

  // setSolid(disk | null) (src/gpu/solver.js:425-440): a RigidDisk from
  // src/cpu/solid.js or any {x, y, r, vx, vy, omega}; its state is sent before
  // every step, as _encodeSolid does.
  setSolid(disk, { passes = 1 } = {}) {
    this.solid = disk || null;
    this._solidPasses = passes;
    if (!disk) addon.setSolid(this._h, null); else this._pushSolid();
  }

  _pushSolid() {
    const d = this.solid;
    if (!d) return;
    addon.setSolid(this._h, { x: d.x, y: d.y, r: d.r, vx: d.vx, vy: d.vy, omega: d.omega,
      passes: this._solidPasses });
  }

  // integrateSolid(dt) (src/gpu/solver.js:454-510): the impulse of the step just
  // taken (FP64 device sum), Newton's third law, the body's own integrate().
  async integrateSolid(dt) {
    const d = this.solid;
    if (!d) return;
    const [fx, fy, tq] = addon.solidImpulse(this._h);
    d.force = [-fx / dt, -fy / dt];
    d.torque = -tq / dt;
    d.integrate(dt);
    const L = this.nx * this.dx;
    d.x = ((d.x % L) + L) % L;
    d.y = ((d.y % L) + L) % L;
  }

  sync() { addon.sync(this._h); }

  setForcing(modes) {
    if (modes === null || modes === undefined || modes.length === 0) { addon.setForcing(this._h, null); return; }
    throw new Error("explicit forcing mode lists are not supported; use setBandForcing({kf, width, amp, drag, seed})");
  }

  setBandForcing(f) { addon.setForcing(this._h, f || null); }

  async maxDivergence() { return addon.maxDivergence(this._h); }

  async poissonResidualAfter(iterations, dt = 0.01) { return addon.poissonResidualAfter(this._h, iterations); }

  stats() { return addon.stats(this._h); }

  diagnostics() { return addon.diagnostics(this._h); }

  destroy() {
    if (this.device && this.device._solvers) this.device._solvers.delete(this);
    addon.destroy(this._h);
  }
}
