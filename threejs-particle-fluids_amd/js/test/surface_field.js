// Surface field of a settling block of liquid through the N-API addon. Writes
// the field's float32 voxels to argv[2] and prints {resolution, voxel} JSON.
const fs = require('fs');
const pf = require('..');

const r = 0.01;
const pos = [];
for (let ix = 0; ix < 5; ix++)
  for (let iy = 0; iy < 4; iy++)
    for (let iz = 0; iz < 5; iz++) pos.push(-0.05 + (ix + 0.5) * 2 * r, 0.01 + iy * 2 * r, -0.05 + (iz + 0.5) * 2 * r);
const n = pos.length / 3;
const ps = new pf.ParticleSystem(n, r);
ps.uploadArrays(new Float32Array(pos));
const fluid = new pf.FluidSystem(ps, {});
const walls = new pf.PrimitiveSet(ps);
walls.addPlane([0, 1, 0], [0, 0, 0]);
const loop = new pf.SimLoop(ps, { substeps: 2, iterations: 3, materials: [fluid], colliders: [walls] });
for (let i = 0; i < 3; i++) loop.step(1 / 60);
const field = new pf.SurfaceField(fluid, {
  bounds: { min: [-0.2, -0.02, -0.2], max: [0.2, 0.2, 0.2] },
  voxelBudget: 100000,
  motionStretch: 0.01,
  wet: true,
});
field.update();
const data = field.readback();
fs.writeFileSync(process.argv[2], Buffer.from(data.buffer));
field.dispose();
console.log(JSON.stringify({ resolution: field.resolution, voxel: field.voxel, particles: n }));
