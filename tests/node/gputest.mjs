// tools/gputest.html checks 1-7 (per-kernel advection and divergence, projection,
// Poisson curves, full steps with red-black and with multigrid MacCormack, scheme
// energies) through the drop-in solver.mjs alone. Writes the GPU results -- arrays as
// raw little-endian files, scalars in results.json -- and the inputs it generated
// (Taylor-Green at the faces, the seeded fields of the reference's JS generator) into
// the directory given as the first argument, for tests/test_gpu_node.py to compare
// with the Python binding on the same inputs.
import { writeFileSync } from "fs";
import { join } from "path";
import { createDevice, GPUFluidSolver } from "../../fluid-sim_amd/node/solver.mjs";

const out = process.argv[2];
const N = 64, dt = 0.01, L = 2 * Math.PI, h = L / N;
const save = (name, arr) => writeFileSync(join(out, name), Buffer.from(arr.buffer, arr.byteOffset, arr.byteLength));

function taylorGreen() {
  const u = new Float64Array(N * N), v = new Float64Array(N * N);
  for (let j = 0; j < N; j++) {
    for (let i = 0; i < N; i++) {
      u[j * N + i] = Math.sin(i * h) * Math.cos((j + 0.5) * h);
      v[j * N + i] = -Math.cos((i + 0.5) * h) * Math.sin(j * h);
    }
  }
  return [u, v];
}

// the generator of tools/gputest.html:113-114, evaluated by the JS engine itself
function seeded(seed) {
  let s = seed;
  const rnd = () => { s = (s * 1103515245 + 12345) & 0x7fffffff; return s / 0x7fffffff - 0.5; };
  const u = new Float64Array(N * N), v = new Float64Array(N * N);
  for (let i = 0; i < u.length; i++) { u[i] = rnd(); v[i] = rnd(); }
  return [u, v];
}

async function main() {
  const { device } = await createDevice();
  const R = {};
  const [tu, tv] = taylorGreen();
  save("ic_tg_u.f64", tu); save("ic_tg_v.f64", tv);

  // 1. advection kernel alone
  {
    const gpu = new GPUFluidSolver(device, { n: N, nu: 0, pressureIterations: 0, advection: "semi-lagrangian" });
    gpu.uploadVelocity(tu, tv);
    gpu.runKernel("advect-sl", dt);
    const { u, v } = await gpu.readVelocity();
    save("adv_u.f32", u); save("adv_v.f32", v);
    gpu.destroy();
  }
  // 2. divergence kernel, read back as rgba texels
  {
    const [u0, v0] = seeded(991);
    save("ic_991_u.f64", u0); save("ic_991_v.f64", v0);
    const gpu = new GPUFluidSolver(device, { n: N, nu: 0 });
    gpu.uploadVelocity(u0, v0);
    gpu.runKernel("divergence");
    const raw = await gpu.readTexture(gpu.div);
    const d = new Float32Array(N * N);
    for (let i = 0; i < N * N; i++) d[i] = raw[i * 4];
    save("div.f32", d);
    let threw = false;
    try { await gpu.readTexture(gpu.output); } catch (e) { threw = true; }
    R.readTextureRejectsHalfFloat = threw;
    gpu.destroy();
  }
  // 3. projection with equal budgets
  {
    const [u0, v0] = seeded(4242);
    save("ic_4242_u.f64", u0); save("ic_4242_v.f64", v0);
    R.projection = {};
    for (const solver of ["jacobi", "red-black"]) {
      const gpu = new GPUFluidSolver(device, { n: N, nu: 0, pressureSolver: solver, pressureIterations: 200 });
      gpu.uploadVelocity(u0, v0);
      const before = await gpu.maxDivergence();
      gpu.step(0.01);
      const after = await gpu.maxDivergence();
      R.projection[solver] = { before, after };
      gpu.destroy();
    }
  }
  // 4. residual curves
  {
    const [u0, v0] = seeded(777);
    save("ic_777_u.f64", u0); save("ic_777_v.f64", v0);
    R.curves = {};
    for (const solver of ["jacobi", "red-black", "multigrid"]) {
      const gpu = new GPUFluidSolver(device, { n: N, nu: 0, pressureSolver: solver });
      gpu.uploadVelocity(u0, v0);
      const budget = solver === "multigrid" ? [0, 1, 2, 3, 4, 5, 6, 8, 10, 12] : [0, 1, 2, 5, 10, 20, 40, 80, 160, 320];
      R.curves[solver] = [];
      for (const it of budget) R.curves[solver].push(await gpu.poissonResidualAfter(it));
      gpu.destroy();
    }
  }
  // 5. full steps, red-black 400 sweeps, semi-Lagrangian
  {
    const gpu = new GPUFluidSolver(device, { n: N, nu: 0, pressureSolver: "red-black", pressureIterations: 400,
      advection: "semi-lagrangian" });
    gpu.uploadVelocity(tu, tv);
    for (let k = 0; k < 10; k++) gpu.step(dt);
    const { u, v } = await gpu.readVelocity();
    save("full_u.f32", u); save("full_v.f32", v);
    R.fullStepMaxDiv = await gpu.maxDivergence();
    gpu.destroy();
  }
  // 6. MacCormack with multigrid, 4 V-cycles
  {
    const gpu = new GPUFluidSolver(device, { n: N, nu: 0, pressureSolver: "multigrid", pressureIterations: 4,
      advection: "maccormack" });
    gpu.uploadVelocity(tu, tv);
    for (let k = 0; k < 10; k++) gpu.step(dt);
    await device.queue.onSubmittedWorkDone();
    const { u, v } = await gpu.readVelocity();
    save("mc_u.f32", u); save("mc_v.f32", v);
    gpu.destroy();
  }
  // 7. kinetic energy after 200 steps per scheme
  R.schemeEnergy = {};
  for (const scheme of ["semi-lagrangian", "maccormack"]) {
    const gpu = new GPUFluidSolver(device, { n: N, nu: 0, pressureSolver: "multigrid", pressureIterations: 3,
      advection: scheme });
    gpu.uploadVelocity(tu, tv);
    for (let k = 0; k < 200; k++) gpu.step(0.01);
    const vel = await gpu.readVelocity();
    let e = 0;
    for (let i = 0; i < vel.u.length; i++) e += vel.u[i] ** 2 + vel.v[i] ** 2;
    R.schemeEnergy[scheme] = 0.5 * e / vel.u.length;
    gpu.destroy();
  }
  // the spectrum and the pressure through the shim's readers
  {
    const gpu = new GPUFluidSolver(device, { n: N, nu: 0.01, pressureSolver: "multigrid", pressureIterations: 20,
      tolerance: 1e-6 });
    gpu.uploadVelocity(tu, tv);
    gpu.step(dt);
    const p = await gpu.readField("pressure");
    save("tg_p.f64", p);
    const { E, counts } = await gpu.energySpectrum();
    save("tg_E.f64", E);
    R.spectrumModes = Array.from(counts).reduce((a, b) => a + b, 0);
    gpu.destroy();
  }
  // splats (each with its own parameters) and the dye through the shim
  {
    const gpu = new GPUFluidSolver(device, { n: N, nu: 0, pressureSolver: "multigrid", pressureIterations: 3,
      dyeFade: 0.02 });
    gpu.uploadVelocity(tu, tv);
    const rgba = new Float32Array(N * N * 4);
    for (let k = 0; k < N * N; k++) { rgba[k * 4] = (k % 7) / 7; rgba[k * 4 + 2] = 0.5; }
    gpu.uploadDye(rgba);
    const splats = [{ x: 1.0, y: 2.0, dx: 3.0, dy: -1.0, radius: 0.3 },
      { x: 6.1, y: 0.1, r: 0.9, g: 0.2, b: 0.5, radius: 0.5, isDye: true }];
    for (let k = 0; k < 3; k++) gpu.step(dt, { splats });
    const raw = await gpu.readTexture(gpu.dyeCur);
    save("dye.f32", raw);
    const { u, v } = await gpu.readVelocity();
    save("splat_u.f32", u); save("splat_v.f32", v);
    // render() as index.html:250-251 calls it; its half-float output refuses readTexture
    const img = gpu.render({ mode: 0, scale: 1.5, bloomThreshold: 0.62, bloomPasses: 2 });
    let threw = false;
    try { await gpu.readTexture(img); } catch (e) { threw = true; }
    R.renderOutputRejected = threw && img === gpu.output;
    save("render.u16", await gpu.readField("output"));
    save("bloom.u16", await gpu.readField("bloom"));
    gpu.destroy();
  }
  // vorticity confinement from the constructor, then through the setter (index.html:471)
  {
    const gpu = new GPUFluidSolver(device, { n: N, nu: 0, pressureSolver: "multigrid", pressureIterations: 3,
      confinement: 1.2 });
    gpu.uploadVelocity(tu, tv);
    for (let k = 0; k < 3; k++) gpu.step(dt);
    gpu.confinement = 2.5;
    for (let k = 0; k < 2; k++) gpu.step(dt);
    const { u, v } = await gpu.readVelocity();
    save("conf_u.f32", u); save("conf_v.f32", v);
    R.confinement = gpu.confinement;
    gpu.destroy();
  }
  writeFileSync(join(out, "results.json"), JSON.stringify(R));
  console.log(JSON.stringify({ ok: true }));
}

main().catch(e => { console.error(e); process.exit(1); });
