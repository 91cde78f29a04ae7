// Node-side check of the learned guess through the drop-in: a 512^2 field built
// from integer residues (the same doubles in any language), the shipped weights
// (at quarter resolution unless argv[2] names another guess mode), five CFL steps
// to the FP64 residual 1e-6; prints a state summary as one JSON line.
import { readFileSync } from "fs";
import { fileURLToPath } from "url";
import { dirname, join } from "path";
import { createDevice, GPUFluidSolver } from "./solver.mjs";

const here = dirname(fileURLToPath(import.meta.url));

async function main() {
  const { device } = await createDevice();
  const n = 512;
  const u = new Float64Array(n * n), v = new Float64Array(n * n);
  for (let j = 0; j < n; j++) {
    for (let i = 0; i < n; i++) {
      u[j * n + i] = ((7 * i + 13 * j) % 31) / 31 - 0.5;
      v[j * n + i] = ((11 * i + 5 * j) % 29) / 29 - 0.5;
    }
  }
  const gpu = new GPUFluidSolver(device, {
    n, nu: 1e-4, advection: "maccormack", pressureSolver: "multigrid",
    pressureIterations: 60, tolerance: 1e-6, mgMinSize: 4,
  });
  gpu.loadGuess(readFileSync(join(here, "..", "weights", "guess_unet_256.fsgn")));
  gpu.setGuess(process.argv[2] || "learned-quarter");
  gpu.uploadVelocity(u, v);
  for (let k = 0; k < 5; k++) gpu.stepCFL(0.5);
  await device.queue.onSubmittedWorkDone();
  const vel = await gpu.readVelocity();
  let su = 0, sv = 0;
  for (let k = 0; k < n * n; k++) { su += vel.u[k] * (k % 97 + 1); sv += vel.v[k] * (k % 89 + 1); }
  const st = gpu.stats();
  console.log(JSON.stringify({ checksumU: su, checksumV: sv, lastResidual: st.lastResidual,
                               totalVCycles: st.totalVCycles }));
  gpu.destroy();
}

main().catch((e) => { console.error(e.stack || e); process.exit(1); });
