// Node-side smoke of the drop-in: Taylor-Green 64^2, MacCormack + multigrid to an
// FP64 residual of 1e-6, five steps; prints the state summary as one JSON line.
import { createDevice, GPUFluidSolver } from "./solver.mjs";

async function main() {
  const { device, info } = await createDevice();
  const n = 64, L = 2 * Math.PI, h = L / n;
  const u = new Float64Array(n * n), v = new Float64Array(n * n);
  for (let j = 0; j < n; j++) {
    for (let i = 0; i < n; i++) {
      u[j * n + i] = Math.sin(i * h) * Math.cos((j + 0.5) * h);
      v[j * n + i] = -Math.cos((i + 0.5) * h) * Math.sin(j * h);
    }
  }
  const gpu = new GPUFluidSolver(device, {
    n, nu: 0.01, advection: "maccormack", pressureSolver: "multigrid",
    pressureIterations: 20, tolerance: 1e-6,
  });
  gpu.uploadVelocity(u, v);
  for (let k = 0; k < 5; k++) gpu.step(0.01);
  await device.queue.onSubmittedWorkDone();
  const vel = await gpu.readVelocity();
  let su = 0, sv = 0;
  for (let k = 0; k < n * n; k++) { su += vel.u[k] * (k % 97 + 1); sv += vel.v[k] * (k % 89 + 1); }
  const st = gpu.stats();
  let threw = false;
  try { new GPUFluidSolver(device, { n, pressureSolver: "sor" }); } catch (e) { threw = true; }
  // a free disk (density 2.5, the RigidDisk update of src/cpu/solid.js) pushed by the
  // same flow for five steps through setSolid / integrateSolid
  const mass = 2.5 * Math.PI * 0.6 * 0.6;
  const disk = {
    x: 2.0, y: 3.5, r: 0.6, vx: 0.1, vy: 0, omega: 0, force: [0, 0], torque: 0,
    mass, inertia: 0.5 * mass * 0.6 * 0.6,
    integrate(dt) {
      this.vx += (this.force[0] / this.mass) * dt;
      this.vy += (this.force[1] / this.mass) * dt;
      this.omega += (this.torque / this.inertia) * dt;
      this.x += this.vx * dt;
      this.y += this.vy * dt;
    },
  };
  const g2 = new GPUFluidSolver(device, {
    n, nu: 0.01, advection: "maccormack", pressureSolver: "multigrid",
    pressureIterations: 20, tolerance: 1e-6,
  });
  g2.uploadVelocity(u, v);
  g2.setSolid(disk, { passes: 2 });
  for (let k = 0; k < 5; k++) { g2.step(0.01); await g2.integrateSolid(0.01); }
  const vel2 = await g2.readVelocity();
  let du = 0;
  for (let k = 0; k < n * n; k++) du += vel2.u[k] * (k % 97 + 1);
  g2.destroy();
  console.log(JSON.stringify({
    device: info.description, checksumU: su, checksumV: sv, t: st.t, steps: st.steps,
    lastResidual: st.lastResidual, maxDivergence: await gpu.maxDivergence(), unknownSelectorThrows: threw,
    disk: { checksumU: du, x: disk.x, y: disk.y, vx: disk.vx, vy: disk.vy, omega: disk.omega },
  }));
  gpu.destroy();
}

main().catch((e) => { console.error(e.stack || e); process.exit(1); });
