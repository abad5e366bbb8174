'use strict';
// JS facade of the native solver, mirroring threejs-particle-fluids' classes
// (src/core/particles.ts, hashGrid/HashGrid.ts, collision/PrimitiveSet.ts,
// loop.ts, src/fluids/FluidSystem.ts, src/softbody/SoftbodySystem.ts,
// src/simulation/Simulation.ts) over the N-API addon. Vectors may be
// three.js-like {x, y, z} objects or [x, y, z] arrays; boxes {min, max}.
// GPU work is enqueued on the library's HIP stream; step() and readbacks
// return Promises like the reference's.

const path = require('path');

const native = require(path.join(__dirname, 'pf_node.node'));

const vec = (v) => (Array.isArray(v) ? [v[0], v[1], v[2]] : [v.x, v.y, v.z]);
const quat = (q) => (q ? (Array.isArray(q) ? [q[0], q[1], q[2], q[3]] : [q.x, q.y, q.z, q.w]) : undefined);
const def = (v, d) => (v === undefined || v === null ? d : v);
const box6 = (b) => [...vec(b.min), ...vec(b.max)];

class Context {
  constructor(device = 0) {
    this.handle = native.contextCreate(device);
  }
  synchronize() {
    native.contextSynchronize(this.handle);
  }
  dispose() {
    if (this.handle) native.contextDestroy(this.handle);
    this.handle = null;
  }
}

let defaultContext = null;
function getDefaultContext() {
  if (!defaultContext) defaultContext = new Context(0);
  return defaultContext;
}

class ParticleSystem {
  /** options.evictable: keep the whole scene in one device arena so a Server can evict it. */
  constructor(capacity, particleRadius, context = getDefaultContext(), options = {}) {
    this.context = context;
    this.handle = native.particlesCreate(context.handle, capacity, particleRadius, !!options.evictable);
    this.capacity = capacity;
    this.particleRadius = particleRadius;
    this.isDisposed = false;
  }
  /** uploadParticles(data: ParticleInit[], start = 0) (particles.ts:121-160). */
  uploadParticles(data, start = 0) {
    const n = data.length;
    const pos = new Float32Array(3 * n);
    const vel = new Float32Array(3 * n);
    const inv = new Float32Array(n);
    const grp = new Uint32Array(n);
    data.forEach((p, i) => {
      pos.set(p.position, 3 * i);
      if (p.velocity) vel.set(p.velocity, 3 * i);
      inv[i] = def(p.invMass, 1);
      grp[i] = def(p.collisionGroup, 0);
    });
    native.particlesUpload(this.handle, start, pos, vel, inv, grp);
  }
  /** Bulk form: xyz triples in Float32Arrays. */
  uploadArrays(positions, start = 0, velocities = null, invMass = null, groups = null) {
    native.particlesUpload(this.handle, start, positions, velocities, invMass, groups);
  }
  setInvMass(range, invMass) {
    native.particlesSetInvMass(this.handle, range.start, range.count, invMass);
  }
  setCollisionGroup(range, group) {
    native.particlesSetCollisionGroup(this.handle, range.start, range.count, group);
  }
  async readback() {
    return native.particlesReadback(this.handle);
  }
  /** Device bytes of an evictable scene's arena (0 otherwise). */
  deviceBytes() {
    return native.particlesDeviceBytes(this.handle);
  }
  get disposed() {
    return this.isDisposed;
  }
  dispose() {
    if (this.isDisposed) return;
    this.isDisposed = true;
    native.particlesDestroy(this.handle);
  }
}

class HashGrid {
  constructor(particles, options, handle = null) {
    this.particles = particles;
    this.owned = handle !== null;
    this.handle =
      handle !== null
        ? handle
        : native.gridCreate(
        particles.handle,
        options.cellSize,
        def(options.hashTableSize, 0),
        options.hashOrigin ? vec(options.hashOrigin) : null,
      );
    const info = native.gridInfo(this.handle);
    this.hashTableSize = info.hashTableSize;
    this.hashTableSizePadded = info.hashTableSizePadded;
    this.cellSize = options ? options.cellSize : undefined;
  }
  async rebuild() {
    native.gridRebuild(this.handle);
  }
  async readback() {
    return native.gridReadback(this.handle);
  }
  async readbackOverflow() {
    return native.gridReadbackOverflow(this.handle);
  }
  setHashOrigin(origin) {
    native.gridSetOrigin(this.handle, vec(origin));
  }
  dispose() {
    if (!this.owned && this.handle) native.gridDestroy(this.handle);
    this.handle = null;
  }
}

const EFFECT = { viscosity: 1, vorticity: 2, surfaceTension: 4, adhesion: 8 };

class FluidSystem {
  constructor(particles, options = {}) {
    this.particles = particles;
    const range = def(options.range, { start: 0, count: particles.capacity });
    this.range = range;
    this.handle = native.fluidCreate(particles.handle, {
      rangeStart: range.start,
      rangeCount: range.count,
      restDensity: options.restDensity,
      particleSpacing: options.particleSpacing,
      smoothingRadius: options.smoothingRadius,
      compliance: options.compliance,
      viscosity: options.viscosity,
      vorticity: options.vorticity,
      surfaceTension: options.surfaceTension,
      adhesion: options.adhesion,
    });
    const info = native.fluidInfo(this.handle);
    this.mass = info.mass;
    this.smoothingRadius = info.smoothingRadius;
    this.particleSpacing = info.particleSpacing;
    this.restDensity = info.restDensity;
    this.neighborRadius = this.smoothingRadius;
  }
  get viscosity() { return native.fluidGet(this.handle, EFFECT.viscosity); }
  set viscosity(v) { native.fluidSet(this.handle, EFFECT.viscosity, v); }
  get vorticity() { return native.fluidGet(this.handle, EFFECT.vorticity); }
  set vorticity(v) { native.fluidSet(this.handle, EFFECT.vorticity, v); }
  get surfaceTension() { return native.fluidGet(this.handle, EFFECT.surfaceTension); }
  set surfaceTension(v) { native.fluidSet(this.handle, EFFECT.surfaceTension, v); }
  get adhesion() { return native.fluidGet(this.handle, EFFECT.adhesion); }
  set adhesion(v) { native.fluidSet(this.handle, EFFECT.adhesion, v); }
  addBoundary(range, options = {}) {
    native.fluidAddBoundary(this.handle, range.start, range.count, def(options.dynamic, true));
  }
  async readbackOverflow() {
    return native.fluidReadbackOverflow(this.handle);
  }
  async readbackDensity() {
    return native.fluidReadbackDensity(this.handle, this.particles.capacity);
  }
  get material() {
    return { kind: 1, handle: this.handle };
  }
  dispose() {
    if (this.handle) native.fluidDestroy(this.handle);
    this.handle = null;
  }
}

/**
 * new ViscositySolver(fluid, {viscosity, iterations = 12}) (src/fluids/ViscositySolver.ts):
 * implicit viscosity for thick liquids. List it directly after its fluid in `materials`.
 */
class ViscositySolver {
  constructor(fluid, options) {
    this.fluid = fluid;
    this.iterations = def(options.iterations, 12);
    this.handle = native.viscosityCreate(fluid.handle, options.viscosity, this.iterations);
    this.value = options.viscosity;
  }
  get viscosity() { return this.value; }
  set viscosity(v) {
    native.viscositySet(this.handle, v);
    this.value = v;
  }
  get material() {
    return { kind: 3, handle: this.handle };
  }
  dispose() {
    if (this.handle) native.viscosityDestroy(this.handle);
    this.handle = null;
  }
}

/**
 * The liquid's surface as a voxel grid (src/fluids/render/SurfaceField.ts:
 * 101-369), the field FluidSurfaceRenderer sphere-traces, computed headless.
 * Options follow FluidSurfaceRendererOptions: bounds ({min, max}),
 * voxelBudget, motionStretch, solids (a particle range); `wet: true` makes
 * the loop's colliders wetted walls. update() enqueues the passes;
 * readback() resolves to a Float32Array of (phi in voxels, volume fraction,
 * wall proximity, 1) per voxel, x fastest.
 */
class SurfaceField {
  constructor(fluid, options) {
    const b = options.bounds;
    const info = native.surfaceFieldCreate(fluid.handle, {
      min: vec(b.min),
      max: vec(b.max),
      voxelBudget: options.voxelBudget,
      motionStretch: options.motionStretch,
      solids: options.solids,
      wet: !!options.wet,
      radius: options.radius,
      meniscusHeight: options.meniscusHeight,
      meniscusWidth: options.meniscusWidth,
    });
    this.fluid = fluid;
    this.handle = info.handle;
    this.resolution = info.resolution;
    this.origin = info.origin;
    this.voxel = info.voxel;
  }
  static defaultVoxelBudget(particleCount) {
    return particleCount >= 50000 ? 1200000 : particleCount >= 20000 ? 900000 : 600000;
  }
  update() {
    native.surfaceFieldUpdate(this.handle);
  }
  readback() {
    const [nx, ny, nz] = this.resolution;
    return native.surfaceFieldReadback(this.handle, nx * ny * nz);
  }
  dispose() {
    if (this.handle) native.surfaceFieldDestroy(this.handle);
    this.handle = null;
  }
}

class PrimitiveSet {
  constructor(particles, options = {}) {
    this.particles = particles;
    this.handle = native.primitivesCreate(particles.handle, def(options.capacity, 0));
  }
  addPlane(normal, point, options = {}) {
    return native.primitivesAdd(this.handle, 'plane', vec(normal), vec(point), options.muS, options.muK);
  }
  addSphere(center, radius, options = {}) {
    return native.primitivesAdd(this.handle, 'sphere', vec(center), radius, !!options.invert, options.muS, options.muK);
  }
  addBox(center, halfExtents, options = {}) {
    return native.primitivesAdd(this.handle, 'box', vec(center), vec(halfExtents), quat(options.rotation),
      !!options.invert, options.muS, options.muK);
  }
  addCapsule(a, b, radius, options = {}) {
    return native.primitivesAdd(this.handle, 'capsule', vec(a), vec(b), radius, options.muS, options.muK);
  }
  setSphere(slot, center, radius, velocity) {
    native.primitivesSetSphere(this.handle, slot, vec(center), radius, velocity ? vec(velocity) : null);
  }
  /** attach()+update() for a pose the caller drives each step. */
  setPose(slot, position, rotation, dt = 0) {
    native.primitivesSetPose(this.handle, slot, vec(position), def(quat(rotation), [0, 0, 0, 1]), dt);
  }
  get count() {
    return native.primitivesCount(this.handle);
  }
  get collider() {
    return { kind: 1, handle: this.handle };
  }
  dispose() {
    if (this.handle) native.primitivesDestroy(this.handle);
    this.handle = null;
  }
}

class SoftbodySystem {
  constructor(particles, options) {
    const mode = def(options.shapeMatching, 'global');
    if (mode !== 'global' && mode !== 'local') {
      throw new Error("SoftbodySystem: shapeMatching must be 'global' or 'local'");
    }
    this.particles = particles;
    this.shapeMatching = mode;
    this.bodies = options.bodies;
    this.handle = native.softbodyCreate(
      particles.handle,
      options.bodies.map((b) => ({
        start: b.range.start,
        count: b.range.count,
        surfaceCount: b.surfaceCount,
        compliance: def(b.compliance, 0),
        restPositions: b.restPositions,
        edges: b.edges ? Uint32Array.from(b.edges) : undefined,
      })),
      mode === 'local' ? 1 : 0,
    );
    if (options.selfCollision) native.softbodySetSelfCollision(this.handle, true);
  }
  setCompliance(index, compliance) {
    native.softbodySetCompliance(this.handle, index, compliance);
  }
  async readbackFrames() {
    return native.softbodyReadbackFrames(this.handle, this.bodies.length);
  }
  get material() {
    return { kind: 2, handle: this.handle };
  }
  dispose() {
    if (this.handle) native.softbodyDestroy(this.handle);
    this.handle = null;
  }
}

/**
 * SDFCollider (SDFCollider.ts:68-311): sdf = {data: Float32Array (x fastest),
 * resolution: [nx, ny, nz], origin: [x, y, z], voxelSize: [x, y, z]}.
 */
class SDFCollider {
  constructor(particles, sdf, options = {}) {
    this.particles = particles;
    this.handle = native.sdfCreate(particles.handle, Float32Array.from(sdf.data), Array.from(sdf.resolution),
      Array.from(sdf.origin), Array.from(sdf.voxelSize), {
        muS: options.muS,
        muK: options.muK,
        position: options.position ? vec(options.position) : undefined,
        rotation: options.rotation ? quat(options.rotation) : undefined,
        scale: options.scale,
        thickness: options.thickness,
      });
  }
  get collider() {
    return { kind: 2, handle: this.handle };
  }
  setPosition(p) { native.sdfSet(this.handle, 'position', vec(p)); }
  setRotation(q) { native.sdfSet(this.handle, 'rotation', quat(q)); }
  setScale(s) { native.sdfSet(this.handle, 'scale', s); }
  setFriction(muS, muK) { native.sdfSet(this.handle, 'friction', muS, muK); }
  set thickness(t) { native.sdfSet(this.handle, 'thickness', t); }
  dispose() {
    if (this.handle) native.sdfDestroy(this.handle);
    this.handle = null;
  }
}

/** SimLoop.contacts (src/core/contact/ContactBuffer.ts:38-98). */
class ContactBuffer {
  constructor(loop, maxContacts) {
    this.loop = loop;
    this.maxContacts = maxContacts;
  }
  /** Pairs emitted in the last substep, dropped ones included. */
  async readbackCount() {
    return native.loopReadbackContacts(this.loop.handle, 0).count;
  }
  /** The stored (i, j) pairs, flattened. */
  async readbackPairs() {
    const r = native.loopReadbackContacts(this.loop.handle, this.maxContacts);
    return { nContacts: r.pairs.length / 2, pairs: r.pairs };
  }
}

/** A material's noSelfContacts range (materials.ts:45): either a native
 * material, or a plain one whose build() returns only noSelfContacts. */
function quietRange(m) {
  if (m.material) return null;
  const built = typeof m.build === 'function' ? m.build({}) : m;
  const keys = Object.keys(built || {}).filter((k) => built[k] !== undefined);
  if (!built || !built.noSelfContacts || keys.some((k) => k !== 'noSelfContacts')) {
    throw new Error('SimLoop: materials other than the built-in ones may only declare noSelfContacts here');
  }
  return built.noSelfContacts;
}

class SimLoop {
  constructor(particles, options = {}) {
    this.particles = particles;
    const materials = def(options.materials, []);
    const quiet = [];
    for (const m of materials) {
      const r = quietRange(m);
      if (r) quiet.push(r.start, r.count);
    }
    const c = options.contact === true ? {} : options.contact || undefined;
    if (c && c.maxContacts !== undefined && !(Number.isInteger(c.maxContacts) && c.maxContacts > 0)) {
      throw new Error(`ContactBuffer: maxContacts must be a positive integer, got ${c.maxContacts}`);
    }
    this.handle = native.loopCreate(
      particles.handle,
      {
        substeps: options.substeps,
        iterations: options.iterations,
        gravity: options.gravity ? vec(options.gravity) : undefined,
        hashOrigin: options.hashOrigin ? vec(options.hashOrigin) : undefined,
        contact: c ? { muS: c.muS, muK: c.muK, maxContacts: c.maxContacts } : undefined,
        quiet: Uint32Array.from(quiet),
      },
      materials.filter((m) => m.material).map((m) => m.material),
      (def(options.colliders, [])).map((col) => col.collider),
    );
    this.contacts = c ? new ContactBuffer(this, def(c.maxContacts, 8 * particles.capacity)) : undefined;
    this.stepping = null;
  }
  /** step(dt) (loop.ts:267-279): resolves when the GPU has finished the step. */
  step(dt) {
    return native.loopStepAsync(this.handle, dt);
  }
  get substeps() { return native.loopInfo(this.handle).substeps; }
  set substeps(v) { native.loopSetSubsteps(this.handle, v); }
  get iterations() { return native.loopInfo(this.handle).iterations; }
  setGravity(g) {
    native.loopSetGravity(this.handle, vec(g));
  }
  useGraph(enabled = true) {
    native.loopSetGraph(this.handle, !!enabled);
  }
  /** Times each tile-union LDS capacity on this loop's workload and keeps the fastest. */
  autotune(dt, candidates = TILE_UNION_CANDIDATES, steps = 2) {
    return native.loopAutotune(this.handle, dt, candidates, steps);
  }
  /** 0: one workgroup per tile; else the persistent form, at most that many
   * workgroups (a multiple of 8) per launch. Results do not depend on it. */
  setLaunchShape(workgroups) {
    native.loopSetLaunchShape(this.handle, workgroups);
  }
  /** Times each launch shape on this loop's workload and keeps the fastest. */
  autotuneLaunch(dt, candidates = LAUNCH_SHAPE_CANDIDATES, steps = 2) {
    return native.loopAutotuneLaunch(this.handle, dt, candidates, steps);
  }
  get hashGrid() {
    return new HashGrid(this.particles, null, native.loopGrid(this.handle));
  }
  async readbackOverflow() {
    return native.loopReadbackOverflow(this.handle);
  }
  dispose() {
    if (this.handle) native.loopDestroy(this.handle);
    this.handle = null;
  }
}

const TILE_UNION_CANDIDATES = [768, 1024, 1280, 1536, 2048];
// 0: one workgroup per tile; then 1, 2, 4 and 8 persistent workgroups per CU of 256.
const LAUNCH_SHAPE_CANDIDATES = [0, 256, 512, 1024, 2048];

/** Many independent SimLoops advanced by one launch sequence (no reference counterpart). */
class Batch {
  constructor(loops, context = getDefaultContext()) {
    this.context = context;
    this.loops = loops;
    this.handle = native.batchCreate(context.handle, loops.map((l) => l.handle));
  }
  /** Resolves when every scene's step has run. */
  step(dt) {
    return native.batchStepAsync(this.context.handle, this.handle, dt);
  }
  useGraph(enabled = true) {
    native.batchSetGraph(this.handle, !!enabled);
  }
  autotune(dt, candidates = TILE_UNION_CANDIDATES, steps = 2) {
    return native.batchAutotune(this.handle, dt, candidates, steps);
  }
  setLaunchShape(workgroups) {
    native.batchSetLaunchShape(this.handle, workgroups);
  }
  autotuneLaunch(dt, candidates = LAUNCH_SHAPE_CANDIDATES, steps = 2) {
    return native.batchAutotuneLaunch(this.handle, dt, candidates, steps);
  }
  dispose() {
    if (this.handle) native.batchDestroy(this.handle);
    this.handle = null;
  }
}

/**
 * The served form (no reference counterpart): step requests for live scenes
 * queue as they arrive and run in multi-scene launches as slots and memory
 * free up; scenes stay resident under an HBM byte budget and idle scenes are
 * evicted to host memory. Scenes need particles created with {evictable: true}.
 */
class Server {
  constructor(options = {}, context = getDefaultContext()) {
    this.context = context;
    this.handle = native.serverCreate(context.handle, options);
  }
  addScene(loop) {
    return native.serverAddScene(this.handle, loop.handle);
  }
  removeScene(id) {
    native.serverRemoveScene(this.handle, id);
  }
  /** Queues one step(dt); returns a Promise of the request's times and status. */
  step(id, dt) {
    const ticket = native.serverSubmit(this.handle, id, dt);
    return native.serverWait(this.handle, ticket);
  }
  submit(id, dt) {
    return native.serverSubmit(this.handle, id, dt);
  }
  wait(ticket) {
    return native.serverWait(this.handle, ticket);
  }
  drain() {
    return native.serverDrain(this.handle);
  }
  evictScene(id) {
    native.serverEvictScene(this.handle, id);
  }
  /** Launch shape of the launches formed from now on (SimLoop.setLaunchShape). */
  setLaunchShape(workgroups) {
    native.serverSetLaunchShape(this.handle, workgroups);
  }
  get stats() {
    return native.serverStats(this.handle);
  }
  dispose() {
    if (this.handle) native.serverDestroy(this.handle);
    this.handle = null;
  }
}

/** new Simulation({particles | particleRadius, container, closed, gravity, substeps, iterations}). */
class Simulation {
  constructor(options) {
    const { particles, particleRadius } = options;
    if ((particles === undefined) === (particleRadius === undefined)) {
      throw new Error('Simulation: give exactly one of `particles` or `particleRadius`');
    }
    if (particles !== undefined && !(Number.isInteger(particles) && particles > 0)) {
      throw new Error(`Simulation: particles must be a positive integer, got ${particles}`);
    }
    this.options = options;
    this.radius = def(particleRadius, 0);
    this.gravityValue = options.gravity ? vec(options.gravity) : [0, -9.81, 0];
    this.fluids = [];
    this.colliderDefs = [];
    this.started = null;
    this.stepping = null;
  }
  addFluid(options) {
    if (this.started) throw new Error('Simulation: addFluid must happen before the first step() or start()');
    if (!options.box) throw new Error('addFluid: give either `box` or `mesh`');
    const entry = {
      box: box6(options.box),
      viscosity: def(options.viscosity, 0.01),
      surfaceTension: def(options.surfaceTension, 0.1),
      vorticity: def(options.vorticity, 0.02),
      adhesion: def(options.adhesion, 0.1),
      thickness: options.thickness,
    };
    this.fluids.push(entry);
    return entry;
  }
  addFloor(options = {}) {
    this.colliderDefs.push({ kind: 'floor', height: def(options.height, 0), friction: def(options.friction, 0.5) });
  }
  addSphere(options) {
    this.colliderDefs.push({ kind: 'sphere', center: vec(def(options.center, [0, 0, 0])), radius: options.radius,
      friction: def(options.friction, 0.5) });
  }
  addBox(options) {
    const s = vec(options.size);
    this.colliderDefs.push({ kind: 'box', center: vec(def(options.center, [0, 0, 0])), half: s.map((v) => v * 0.5),
      rotation: quat(options.rotation), friction: def(options.friction, 0.5) });
  }
  addCapsule(options) {
    this.colliderDefs.push({ kind: 'capsule', a: vec(options.start), b: vec(options.end), radius: options.radius,
      friction: def(options.friction, 0.5) });
  }
  get gravity() {
    return this.gravityValue;
  }
  get particleRadius() {
    return this.radius;
  }
  get particleCount() {
    return this.system ? this.system.capacity : 0;
  }
  get particleSystem() {
    if (!this.system) throw new Error('Simulation.particleSystem is created on the first step');
    return this.system;
  }
  get loop() {
    if (!this.simLoop) throw new Error('Simulation.loop is created on the first step');
    return this.simLoop;
  }
  start() {
    if (!this.started) {
      this.started = Promise.resolve().then(() => this.build()).catch((error) => {
        this.started = null;
        throw error;
      });
    }
    return this.started;
  }
  step(dt = 1 / 60) {
    if (!this.stepping) {
      this.stepping = this.start()
        .then(() => {
          this.simLoop.setGravity(this.gravityValue);
          return this.simLoop.step(dt);
        })
        .finally(() => {
          this.stepping = null;
        });
    }
    return this.stepping;
  }
  build() {
    if (!this.fluids.length) throw new Error('Simulation: add a fluid, smoke, soft body, or cloth before stepping');
    if (this.colliderDefs.some((d) => d.kind !== 'floor') && this.options.particles !== undefined) {
      // obstacles change the fitted layout; the Python host implements clearOf for them
      throw new Error('Simulation: budget fitting with obstacles is not available in the JS facade');
    }
    const container = this.options.container ? box6(this.options.container) : null;
    let r = this.radius;
    if (this.options.particles !== undefined) {
      const fit = native.layoutFitBudget(this.options.particles, container, this.fluids.flatMap((f) => f.box));
      r = fit.radius;
    }
    this.radius = r;
    const parts = [];
    const ranges = [];
    let total = 0;
    for (const f of this.fluids) {
      let b = f.box;
      if (container) {
        b = [Math.max(b[0], container[0] + r), Math.max(b[1], container[1] + r), Math.max(b[2], container[2] + r),
          Math.min(b[3], container[3] - r), Math.min(b[4], container[4] - r), Math.min(b[5], container[5] - r)];
      }
      const pts = native.layoutFillBox(b.slice(0, 3), b.slice(3), 2 * r);
      if (pts.length === 0) throw new Error('addFluid: the fluid has no room; check its box and the container');
      parts.push(pts);
      ranges.push({ start: total, count: pts.length / 3 });
      total += pts.length / 3;
    }
    const positions = new Float32Array(3 * total);
    let off = 0;
    for (const p of parts) {
      positions.set(p, off);
      off += p.length;
    }
    this.system = new ParticleSystem(total, r);
    this.system.uploadArrays(positions, 0);
    this.fluidSystems = this.fluids.map((f, i) => new FluidSystem(this.system, {
      range: ranges[i], viscosity: f.viscosity, surfaceTension: f.surfaceTension, vorticity: f.vorticity,
      adhesion: f.adhesion,
    }));
    // thickness adds a ViscositySolver right after its fluid (Simulation.ts:740-745)
    this.materials = [];
    this.viscositySolvers = [];
    this.fluids.forEach((f, i) => {
      this.materials.push(this.fluidSystems[i]);
      if (f.thickness !== undefined && f.thickness !== null) {
        const solver = new ViscositySolver(this.fluidSystems[i], {
          viscosity: f.thickness, iterations: Math.round(16 * Math.cbrt(ranges[i].count / 10000)),
        });
        this.viscositySolvers.push(solver);
        this.materials.push(solver);
      }
    });
    this.walls = new PrimitiveSet(this.system);
    if (container) {
      const mn = container.slice(0, 3);
      const mx = container.slice(3);
      this.walls.addPlane([0, 1, 0], mn);
      this.walls.addPlane([1, 0, 0], mn);
      this.walls.addPlane([-1, 0, 0], mx);
      this.walls.addPlane([0, 0, 1], mn);
      this.walls.addPlane([0, 0, -1], mx);
      if (this.options.closed) this.walls.addPlane([0, -1, 0], mx);
    }
    for (const d of this.colliderDefs) {
      const muS = d.friction;
      const muK = d.friction * 0.8;
      if (d.kind === 'floor') this.walls.addPlane([0, 1, 0], [0, d.height, 0], { muS, muK });
      if (d.kind === 'sphere') this.walls.addSphere(d.center, d.radius, { muS, muK });
      if (d.kind === 'box') this.walls.addBox(d.center, d.half, { rotation: d.rotation, muS, muK });
      if (d.kind === 'capsule') this.walls.addCapsule(d.a, d.b, d.radius, { muS, muK });
    }
    const substeps = def(this.options.substeps, Math.min(24, Math.ceil(4 * Math.max(1, 0.018 / r))));
    this.simLoop = new SimLoop(this.system, {
      substeps,
      iterations: def(this.options.iterations, 2),
      gravity: this.gravityValue,
      materials: this.materials,
      colliders: this.walls.count > 0 ? [this.walls] : [],
    });
  }
  dispose() {
    if (this.simLoop) this.simLoop.dispose();
    for (const v of this.viscositySolvers || []) v.dispose();
    for (const f of this.fluidSystems || []) f.dispose();
    if (this.walls) this.walls.dispose();
    if (this.system) this.system.dispose();
  }
}

// ---------------------------------------------------------------- formats
/** decodeSdfBinary(buffer) (src/sdf/writeBinary.ts:54-106): SDFData from PSDF bytes. */
function decodeSdfBinary(buffer) {
  const bytes = buffer instanceof Uint8Array ? buffer : new Uint8Array(buffer.buffer ? buffer.buffer : buffer);
  return native.sdfDecode(bytes);
}
/** encodeSdfBinary(sdf) (src/sdf/writeBinary.ts:25-52): PSDF bytes (an ArrayBuffer). */
function encodeSdfBinary(sdf) {
  const data = sdf.data instanceof Float32Array ? sdf.data : Float32Array.from(sdf.data);
  const out = native.sdfEncode(data, Array.from(sdf.resolution), Array.from(sdf.origin), Array.from(sdf.voxelSize));
  return out.buffer;
}
/**
 * A baked soft-body template (SampledBody, demo/presets/elastic.ts:26-34):
 * positions, per-particle surface flags (surface first), local edges. Returns
 * the pieces a SoftbodySystem body takes.
 */
function loadSampledBody(json) {
  const body = typeof json === 'string' ? JSON.parse(json) : json;
  if (body.positions.length % 3) throw new Error('SampledBody: positions must hold xyz triples');
  const count = body.positions.length / 3;
  const surface = body.surface || new Array(count).fill(1);
  if (surface.length !== count) throw new Error(`SampledBody: ${surface.length} surface flags for ${count} particles`);
  const surfaceCount = surface.filter(Boolean).length;
  for (let i = 0; i < count; i++) {
    if ((i < surfaceCount) !== !!surface[i]) throw new Error('SampledBody: surface particles must come first');
  }
  const edges = Uint32Array.from(body.edges || []);
  if (edges.length % 2) throw new Error('SampledBody: edges must be index pairs');
  for (let e = 0; e < edges.length; e += 2) {
    if (!(edges[e] < edges[e + 1] && edges[e + 1] < count)) {
      throw new Error('SampledBody: edges must be pairs a < b of particle indices');
    }
  }
  return { positions: Float32Array.from(body.positions), surfaceCount, edges, scale: def(body.scale, 1) };
}
/** The demo's deterministic [0, 1) generator (demo/presets/shared.ts:66-74): mulberry32. */
function seededRandom(seed) {
  let state = seed | 0;
  return () => {
    state = (state + 0x6d2b79f5) | 0;
    let t = Math.imul(state ^ (state >>> 15), 1 | state);
    t = (t + Math.imul(t ^ (t >>> 7), 61 | t)) ^ t;
    return ((t ^ (t >>> 14)) >>> 0) / 4294967296;
  };
}

module.exports = {
  native,
  Context,
  ParticleSystem,
  HashGrid,
  FluidSystem,
  PrimitiveSet,
  SDFCollider,
  SoftbodySystem,
  SimLoop,
  ContactBuffer,
  SurfaceField,
  decodeSdfBinary,
  encodeSdfBinary,
  loadSampledBody,
  seededRandom,
  Simulation,
  ViscositySolver,
  Batch,
  Server,
  getDefaultContext,
};
