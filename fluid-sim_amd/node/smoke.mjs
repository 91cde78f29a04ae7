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
  console.log(JSON.stringify({
    device: info.description, checksumU: su, checksumV: sv, t: st.t, steps: st.steps,
    lastResidual: st.lastResidual, maxDivergence: await gpu.maxDivergence(), unknownSelectorThrows: threw,
  }));
  gpu.destroy();
}

main().catch((e) => { console.error(e.stack || e); process.exit(1); });
