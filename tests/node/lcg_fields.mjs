// The seeded fields of tools/gputest.html:113-114 (s = (s * 1103515245 + 12345) &
// 0x7fffffff in IEEE doubles, then s / 0x7fffffff - 0.5), evaluated by the JS engine:
// writes u then v (n*n doubles each) for each seed given, to the file named first.
import { writeFileSync } from "fs";

const [file, n, ...seeds] = process.argv.slice(2);
const N = Number(n);
const parts = [];
for (const seed of seeds) {
  let s = Number(seed);
  const rnd = () => { s = (s * 1103515245 + 12345) & 0x7fffffff; return s / 0x7fffffff - 0.5; };
  const u = new Float64Array(N * N), v = new Float64Array(N * N);
  for (let i = 0; i < u.length; i++) { u[i] = rnd(); v[i] = rnd(); }
  parts.push(Buffer.from(u.buffer), Buffer.from(v.buffer));
}
writeFileSync(file, Buffer.concat(parts));
