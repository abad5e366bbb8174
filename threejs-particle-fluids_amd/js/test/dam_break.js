'use strict';
// Dam break through the JS facade (Simulation -> addFluid -> step()), as
// examples/fluid.ts builds it; writes the final positions (vec4 f32) to argv[2].
const fs = require('fs');
const path = require('path');

const pf = require(path.join(__dirname, '..', 'index.js'));

async function main() {
  const out = process.argv[2];
  const budget = Number(process.argv[3] || 2000);
  const steps = Number(process.argv[4] || 5);
  const thickness = process.argv[5] === undefined ? undefined : Number(process.argv[5]);
  const sim = new pf.Simulation({
    particles: budget,
    container: { min: [-0.5, 0, -0.3], max: [0.5, 0.8, 0.3] },
    substeps: 2,
    iterations: 4,
  });
  sim.addFluid({ box: { min: [-0.5, 0, -0.3], max: [-0.1, 0.5, 0.3] }, thickness });
  for (let i = 0; i < steps; i++) await sim.step(1 / 60);
  const snap = await sim.particleSystem.readback();
  fs.writeFileSync(out, Buffer.from(snap.positions.buffer));
  const overflow = await sim.loop.readbackOverflow();
  console.log(JSON.stringify({ particles: sim.particleCount, radius: sim.particleRadius, overflow }));
  sim.dispose();
}

main().catch((e) => {
  console.error(e.stack || String(e));
  process.exit(1);
});
