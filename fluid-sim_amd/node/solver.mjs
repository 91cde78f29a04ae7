// solver.mjs -- drop-in for the reference's src/gpu/solver.js, backed by the
// MI355X library through the Node-API addon (fluidsim.node -> libfluidsim.so).
//
// Exports the same two names (createDevice, GPUFluidSolver) with the same
// constructor options, defaults, methods and Promise-returning readbacks (the
// copies run on a worker thread; calls on one solver keep their call order).
// Deliberate differences: unknown selectors throw (the reference silently runs
// Jacobi / semi-Lagrangian, src/gpu/solver.js:606, :725); each splat keeps its own
// parameters (the reference binds the first one's, :546-553); render() writes
// `output` and `bloom` as the reference's, but there is no canvas to present them
// (readField("output") returns their half-float bits). setForcing(modes, opts)
// is the reference's mode-list forcing; setBandForcing() adds the CPU solver's
// seeded band forcing (src/cpu/solver.js:179-237). The reference's internal
// textures are named objects here (velCur, velNext, presCur, div): readTexture()
// returns their rgba32float texels, runKernel() replaces gputest's dispatches.
// Written in Node 12 syntax (no optional chaining, no nullish coalescing).
import { createRequire } from "module";

const require = createRequire(import.meta.url);
const addon = require("./fluidsim.node");

const ADVECTION = { "semi-lagrangian": 0, maccormack: 1 };
const PRESSURE = { jacobi: 0, "red-black": 1, multigrid: 2 };
const GUESS = { zero: 0, previous: 1, learned: 2, "learned-coarse": 3, "learned-quarter": 4,
  extrapolated: 5 };
// fs_run_kernel ids (include/fluidsim.h)
const KERNEL = { "advect-sl": 0, "advect-mc": 1, divergence: 2, diffuse: 3, project: 4, gradient: 5, solid: 6,
  forcing: 7, "advect-dye": 8, confine: 9 };
// fs_read_field ids
const FIELD = { u: 0, v: 1, div: 2, pressure: 3, residual: 4, dye: 5, output: 6, bloom: 7 };

export async function createDevice({ requireTimestamps = false } = {}) {
  const d = addon.device();
  if (!d.count) throw new Error("no HIP device available");
  const solvers = new Set();
  const device = {
    _solvers: solvers,
    queue: {
      // the reference awaits this after step() to time the GPU
      onSubmittedWorkDone: async () => {
        for (const s of solvers) await addon.syncAsync(s._h);
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
    mgCoarseSolve = "sweeps",
  } = {}) {
    if (!(advection in ADVECTION)) {
      throw new Error(`unknown advection scheme ${advection}; have ${Object.keys(ADVECTION).join(", ")}`);
    }
    if (!(pressureSolver in PRESSURE)) {
      throw new Error(`unknown pressure solver ${pressureSolver}; have ${Object.keys(PRESSURE).join(", ")}`);
    }
    if (!(guess in GUESS)) throw new Error(`unknown guess mode ${guess}`);
    if (!(dyeFade >= 0 && dyeFade <= 1)) throw new Error("dyeFade must be in [0, 1]");
    if (mgCoarseSolve !== "sweeps" && mgCoarseSolve !== "exact") {
      throw new Error(`unknown coarse solve ${mgCoarseSolve}; have sweeps, exact`);
    }
    this.device = device;
    this.nx = nx; this.ny = ny;
    this.dx = lx / nx; this.dy = ly / ny;
    this._nu = nu;
    this._confinement = confinement;
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
      guess: GUESS[guess], solidMaskWidth, dyeFade, confinement, mgCoarseExact: mgCoarseSolve === "exact" ? 1 : 0,
    });
    if (device && device._solvers) device._solvers.add(this);
    // the reference's textures by name (src/gpu/solver.js:111-125, :287-292), for readTexture
    this.div = { format: "rgba32float", _field: "div" };
    this.output = { format: "rgba16float", _field: "output" };
    // half resolution; after render() bloom[0] holds the last blur pass (the reference
    // reverses the pair after each pass, src/gpu/solver.js:808-821)
    this.bloom = [{ format: "rgba16float", _field: "bloom" }, { format: "rgba16float", _field: null }];
  }

  get velCur() { return { format: "rgba32float", _field: "velocity" }; }
  get velNext() { return { format: "rgba32float", _field: "velocity" }; }
  get presCur() { return { format: "rgba32float", _field: "pressure" }; }
  get presNext() { return { format: "rgba32float", _field: "pressure" }; }
  get dyeCur() { return { format: "rgba32float", _field: "dye" }; }
  get dyeNext() { return { format: "rgba32float", _field: "dye" }; }

  get nu() { return this._nu; }
  set nu(v) { addon.setParam(this._h, "nu", v); this._nu = v; }

  // vorticity confinement (src/gpu/solver.js:558-568), mutable as index.html:471 uses it
  get confinement() { return this._confinement; }
  set confinement(v) { addon.setParam(this._h, "confinement", v); this._confinement = v; }

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

  readVelocity() { return addon.readVelocity(this._h, this.nx * this.ny); }

  // readTexture(tex) (src/gpu/solver.js:336-353): Float32Array(nx*ny*4) of rgba texels,
  // the velocity in (r, g), a scalar field in r; throws on a texture that is not
  // rgba32float, as the reference does
  async readTexture(tex) {
    if (!tex || tex.format !== "rgba32float") {
      throw new Error(`readTexture expects rgba32float, got ${tex ? tex.format : tex}; ` +
        "the 16-byte stride would misread it");
    }
    const n = this.nx * this.ny;
    const out = new Float32Array(n * 4);
    if (tex._field === "velocity") {
      const { u, v } = await addon.readVelocity(this._h, n);
      for (let i = 0; i < n; i++) { out[i * 4] = u[i]; out[i * 4 + 1] = v[i]; }
    } else if (tex._field === "dye") {
      out.set(await addon.readField(this._h, FIELD.dye, n));
    } else {
      const f = await addon.readField(this._h, FIELD[tex._field], n);
      for (let i = 0; i < n; i++) out[i * 4] = f[i];
    }
    return out;
  }

  // a field by name: "u", "v", "div", "pressure" (Float64Array), "residual", "dye"; the
  // render targets "output" and "bloom" as Uint16Array of rgba IEEE half floats
  readField(name) {
    if (!(name in FIELD)) throw new Error(`unknown field ${name}; have ${Object.keys(FIELD).join(", ")}`);
    const count = name === "output" ? 4 * this.nx * this.ny
      : name === "bloom" ? 4 * (this.nx >> 1) * (this.ny >> 1) : this.nx * this.ny;
    return addon.readField(this._h, FIELD[name], count);
  }

  // render({mode, scale, bloomThreshold, bloomPasses}) (src/gpu/solver.js:771-823): the
  // display image of the current velocity and dye into `output`, then the bloom chain;
  // returns `output`. Its pixels are rgba16float, which readTexture refuses as the
  // reference's does; readField("output") / readField("bloom") return their bits.
  render({ mode = 0, scale = 1.5, bloomThreshold = 0.75, bloomPasses = 2 } = {}) {
    let passes = 0;
    for (let k = 0; k < bloomPasses; k++) passes++;
    addon.render(this._h, mode, scale, bloomThreshold, passes);
    return this.output;
  }

  // one kernel by name (replaces gputest's _dispatch of internals, tools/gputest.html:90-129)
  runKernel(name, dt = 0) {
    if (!(name in KERNEL)) throw new Error(`unknown kernel ${name}; have ${Object.keys(KERNEL).join(", ")}`);
    addon.runKernel(this._h, KERNEL[name], dt);
  }

  // energySpectrum (src/core/fft.js:196-234) of the current velocity on the device:
  // Promise<{k, E, counts}> for shells k = 0 .. n/2
  async energySpectrum() {
    const len = Math.floor(this.nx / 2) + 1;
    const { E, counts } = await addon.energySpectrum(this._h, len);
    const k = new Float64Array(len);
    for (let i = 0; i < len; i++) k[i] = i;
    return { k, E, counts };
  }

  // step(dt, {splats}) (src/gpu/solver.js:518-591); each splat keeps its own parameters
  step(dt, { splats = [] } = {}) {
    this._pushSolid();
    if (splats.length) addon.stepSplats(this._h, dt, splats);
    else addon.step(this._h, dt);
  }

  // uploadDye(rgba) (src/gpu/solver.js:317-323): Float32Array(nx*ny*4)
  uploadDye(rgba) { addon.uploadDye(this._h, rgba); }

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

  // setForcing(modes, {amplitude, drag, phaseDrift}) (src/gpu/solver.js:370-414)
  setForcing(modes, { amplitude = 1, drag = 0, phaseDrift = 0.3 } = {}) {
    if (!modes || modes.length === 0) { addon.setForcing(this._h, null); return; }
    addon.setForcingModes(this._h, modes.map(m => m.kx), modes.map(m => m.ky), modes.map(m => m.phase || 0),
      amplitude, drag, phaseDrift);
  }

  setBandForcing(f) { addon.setForcing(this._h, f || null); }

  maxDivergence() { return addon.maxDivergence(this._h); }

  poissonResidualAfter(iterations, dt = 0.01) { return addon.poissonResidualAfter(this._h, iterations); }

  stats() { return addon.stats(this._h); }

  diagnostics() { return addon.diagnostics(this._h); }

  destroy() {
    if (this.device && this.device._solvers) this.device._solvers.delete(this);
    addon.destroy(this._h);
  }
}
