'use strict';
// Particle contacts through the JS API, as the reference's contact tests
// drive them: a 3^3 pile dropped on a plane (contact-non-interpenetration)
// written to argv[2] after argv[3] frames, and the emitter-subset search
// (contact-generate-emitters) with a material that only declares
// noSelfContacts. Prints a JSON summary.
const fs = require('fs');
const path = require('path');

const pf = require(path.join(__dirname, '..', 'index.js'));

async function pile(out, frames) {
  const r = 0.05;
  const sp = 2 * r * 1.05;
  const init = [];
  for (let ix = 0; ix < 3; ix++)
    for (let iy = 0; iy < 3; iy++)
      for (let iz = 0; iz < 3; iz++)
        init.push({ position: [(ix - 1) * sp, 0.5 + (iy + 1) * sp, (iz - 1) * sp], velocity: [0, 0, 0], invMass: 1 });
  const particles = new pf.ParticleSystem(init.length, r);
  particles.uploadParticles(init);
  const floor = new pf.PrimitiveSet(particles);
  floor.addPlane([0, 1, 0], [0, 0, 0]);
  const loop = new pf.SimLoop(particles, {
    substeps: 4,
    iterations: 4,
    contact: { maxContacts: 256, muS: 0.4, muK: 0.3 },
    colliders: [floor],
  });
  for (let n = 0; n < frames; n++) await loop.step(1 / 60);
  const snap = await particles.readback();
  fs.writeFileSync(out, Buffer.from(snap.positions.buffer));
  const count = await loop.contacts.readbackCount();
  loop.dispose();
  floor.dispose();
  particles.dispose();
  return count;
}

async function emitters() {
  const r = 0.025;
  const sp = 2 * r;
  const col = 200;
  const slab = 60;
  const side = Math.ceil(Math.cbrt(col));
  const init = [];
  for (let i = 0; i < col; i++) {
    const x = i % side, y = Math.floor(i / side) % side, z = Math.floor(i / (side * side));
    init.push({ position: [-side * sp * 0.5 + (x + 0.5) * sp, (y + 0.5) * sp, -side * sp * 0.5 + (z + 0.5) * sp],
                invMass: 1, collisionGroup: 1 });
  }
  const ss = Math.ceil(Math.sqrt(slab));
  for (let i = 0; i < slab; i++) {
    const x = i % ss, z = Math.floor(i / ss);
    init.push({ position: [-ss * sp * 0.5 + (x + 0.5) * sp, -sp * 0.5, -ss * sp * 0.5 + (z + 0.5) * sp],
                invMass: 0, collisionGroup: 2 });
  }
  const run = async (quiet) => {
    const particles = new pf.ParticleSystem(init.length, r);
    particles.uploadParticles(init);
    const materials = quiet ? [{ build: () => ({ noSelfContacts: quiet }) }] : [];
    const loop = new pf.SimLoop(particles, { substeps: 1, iterations: 1, contact: { maxContacts: init.length * 8 }, materials });
    await loop.step(1 / 60);
    const { pairs } = await loop.contacts.readbackPairs();
    loop.dispose();
    particles.dispose();
    return pairs;
  };
  const toSet = (pairs) => {
    const s = new Set();
    for (let p = 0; p < pairs.length / 2; p++) s.add(`${Math.min(pairs[2 * p], pairs[2 * p + 1])}-${Math.max(pairs[2 * p], pairs[2 * p + 1])}`);
    return s;
  };
  const all = await run();
  const emit = await run({ start: 0, count: col });
  const a = toSet(all);
  const e = toSet(emit);
  let same = a.size === e.size;
  for (const k of a) same = same && e.has(k);
  let fromSlab = true;
  for (let p = 0; p < emit.length / 2; p++) fromSlab = fromSlab && emit[2 * p] >= col;
  return { pairs: e.size, same, fromSlab };
}

async function main() {
  const count = await pile(process.argv[2], Number(process.argv[3] || 30));
  const em = await emitters();
  console.log(JSON.stringify({ count, ...em }));
}

main().catch((e) => {
  console.error(e.stack || String(e));
  process.exit(1);
});
