'use strict';
// The served form through the JS facade: dam-break scenes of different sizes
// (evictable particles) stepped by a Server, each compared with the same
// scene stepped as a plain SimLoop; prints a JSON summary.
const path = require('path');

const pf = require(path.join(__dirname, '..', 'index.js'));

const CONTAINER = [-0.5, 0, -0.3, 0.5, 0.8, 0.3];
const FLUID = [-0.5, 0, -0.3, -0.1, 0.5, 0.3];

function scene(budget, evictable) {
  const fit = pf.native.layoutFitBudget(budget, CONTAINER, FLUID);
  const r = fit.radius;
  const lo = FLUID.slice(0, 3).map((v, a) => Math.max(v, CONTAINER[a] + r));
  const hi = FLUID.slice(3).map((v, a) => Math.min(v, CONTAINER[3 + a] - r));
  const pos = pf.native.layoutFillBox(lo, hi, 2 * r);
  const n = pos.length / 3;
  const ps = new pf.ParticleSystem(n, r, pf.getDefaultContext(), { evictable });
  ps.uploadArrays(pos);
  const fluid = new pf.FluidSystem(ps, { viscosity: 0.01, surfaceTension: 0.1, vorticity: 0.02, adhesion: 0.1 });
  const walls = new pf.PrimitiveSet(ps);
  walls.addPlane([0, 1, 0], [0, 0, 0]);
  walls.addPlane([1, 0, 0], [-0.5, 0, 0]);
  walls.addPlane([-1, 0, 0], [0.5, 0, 0]);
  walls.addPlane([0, 0, 1], [0, 0, -0.3]);
  walls.addPlane([0, 0, -1], [0, 0, 0.3]);
  const loop = new pf.SimLoop(ps, { substeps: 2, iterations: 4, materials: [fluid], colliders: [walls] });
  return { ps, fluid, walls, loop };
}

async function main() {
  const budgets = [2000, 5000, 3000];
  const steps = 3;
  const plain = budgets.map((b) => scene(b, false));
  const served = budgets.map((b) => scene(b, true));
  const srv = new pf.Server({ slots: 2 });
  const ids = served.map((s) => srv.addScene(s.loop));
  const pending = [];
  for (let k = 0; k < steps; k++) for (const id of ids) pending.push(srv.step(id, 1 / 60));
  const infos = await Promise.all(pending);
  await srv.drain();
  let rejected = false;
  try {
    await served[0].loop.step(1 / 60);
  } catch (e) {
    rejected = /being served/.test(e.message);
  }
  const stats = srv.stats;
  for (const id of ids) srv.removeScene(id);
  for (const s of plain) for (let k = 0; k < steps; k++) await s.loop.step(1 / 60);
  let same = true;
  for (let i = 0; i < budgets.length; i++) {
    const a = await plain[i].ps.readback();
    const b = await served[i].ps.readback();
    for (let j = 0; j < a.positions.length; j++) {
      if (a.positions[j] !== b.positions[j] || a.velocities[j] !== b.velocities[j]) same = false;
    }
  }
  console.log(JSON.stringify({
    same, rejected, done: infos.every((r) => r.status === 'done'), completed: stats.completed,
    maxLaunchScenes: stats.maxLaunchScenes,
  }));
  srv.dispose();
}

main().catch((e) => {
  console.error(e.stack || String(e));
  process.exit(1);
});
