'use strict';
// Test driver for node/gpuworld.js: reads a JSON job (an XpbdWorld-shaped
// world written by the Python tests, +-Infinity as strings), runs it through
// the gpuworld surface and prints one JSON result line.
//   node tests/node_driver.js pure JOB.json   pure helpers only (no GPU)
//   node tests/node_driver.js gpu JOB.json    create/upload/step/download, rollout
const fs = require('fs');
const path = require('path');
const gw = require(path.join(__dirname, '..', 'node', 'gpuworld.js'));

function revive(key, v) {
  if (v === 'Infinity') return Infinity;
  if (v === '-Infinity') return -Infinity;
  if (v === 'NaN') return NaN;
  return v;
}

function enc(x) {
  if (Number.isNaN(x)) return 'NaN';
  if (x === Infinity) return 'Infinity';
  if (x === -Infinity) return '-Infinity';
  return x;
}

function makeWorld(w) {
  const n = w.n;
  const f64 = (a) => Float64Array.from(a);
  const cpu = {
    n, substeps: w.substeps, iterations: w.iterations, h: w.h,
    linDrag: w.linDrag, angDrag: w.angDrag, maxSpeed: w.maxSpeed, maxAngSpeed: w.maxAngSpeed,
    px: f64(w.px), py: f64(w.py), vx: f64(w.vx), vy: f64(w.vy), ang: f64(w.ang), av: f64(w.av),
    ppx: f64(w.px), ppy: f64(w.py), angP: f64(w.ang),
    im: f64(w.im), ii: f64(w.ii), fx: new Float64Array(n), fy: new Float64Array(n), tq: new Float64Array(n),
    constraints: w.constraints.map((c) => Object.assign({}, c)),
    collPts: w.collPts,
    stats: { maxSpeed: 0, maxAngSpeed: 0, nan: 0, collisions: 0, broken: 0 },
  };
  if (w.colorOff) { cpu.colorOff = Uint32Array.from(w.colorOff); cpu.colorCnt = Uint32Array.from(w.colorCnt); }
  return cpu;
}

function state(cpu) {
  const cs = cpu.constraints;
  return {
    px: Array.from(cpu.px.subarray(0, cpu.n)), py: Array.from(cpu.py.subarray(0, cpu.n)),
    vx: Array.from(cpu.vx.subarray(0, cpu.n)), vy: Array.from(cpu.vy.subarray(0, cpu.n)),
    ang: Array.from(cpu.ang.subarray(0, cpu.n)), av: Array.from(cpu.av.subarray(0, cpu.n)),
    lambda: cs.map((c) => enc(c.lambda)), broken: cs.map((c) => (c.broken ? 1 : 0)),
    force: cs.map((c) => enc(c.force || 0)), color: cs.map((c) => c.color),
    owner: cs.map((c) => c.owner), a: cs.map((c) => c.a), b: cs.map((c) => c.b),
  };
}

async function pure(job) {
  const cpu = makeWorld(job.world);
  const cd = new Float32Array(cpu.constraints.length * gw.GPU_STRIDES.constraint);
  gw.packControls(cpu, cd);
  const packed = Array.from(new Uint32Array(cd.buffer));
  const bodyIn = new Float32Array(Uint32Array.from(job.bodyIn).buffer);
  const consIn = new Float32Array(Uint32Array.from(job.consIn).buffer);
  const stats = gw.mirrorInto(cpu, bodyIn, consIn);
  const g1 = gw.groupCollPoints(job.groupBodies.map((b) => ({ body: b })), job.nB);
  const g2 = gw.groupCollPoints(job.shuffledBodies.map((b) => ({ body: b })), job.nB);
  let threw = false;
  try {
    gw.groupCollPoints([{ body: 0 }, { body: 5 }], 3);
  } catch (e) {
    threw = true;
  }
  return {
    packed,
    mirrored: {
      px: Array.from(cpu.px.subarray(0, cpu.n)).map(enc), py: Array.from(cpu.py.subarray(0, cpu.n)).map(enc),
      lambda: cpu.constraints.map((c) => enc(c.lambda)), broken: cpu.constraints.map((c) => (c.broken ? 1 : 0)),
      force: cpu.constraints.map((c) => enc(c.force || 0)), stats: { maxSpeed: enc(stats.maxSpeed),
        maxAngSpeed: enc(stats.maxAngSpeed), nan: stats.nan },
    },
    group: { order: Array.from(g1.order), range: Array.from(g1.range),
             shuffledOrder: Array.from(g2.order), shuffledRange: Array.from(g2.range) },
    groupThrows: threw,
    available: gw.isAvailable(),
  };
}

async function gpu(job) {
  const out = {};
  // host-driven loop: upload once, then per step uploadControls + step, download at the end
  const cpu = makeWorld(job.world);
  const g = await gw.create(cpu, { rects: job.world.rects }, null, null, { path: job.path || 'persistent' });
  gw.upload(g);
  for (let s = 0; s < job.steps; s++) {
    const t = job.targets[s];
    for (let k = 0; k < cpu.constraints.length; k++) {
      const c = cpu.constraints[k];
      if (c.type === gw.C_ANGLE) c.target = t[c.owner % t.length];
    }
    gw.uploadControls(g);
    await gw.step(g);
  }
  const stats = await gw.download(g);
  out.host = Object.assign(state(cpu), { stats, dispatches: g.dispatches });
  gw.destroy(g);
  gw.destroy(g);
  // device-side control law rollout with events
  if (job.law) {
    const cpu2 = makeWorld(job.world);
    const cr = {};
    for (const [k, v] of Object.entries(job.creatures)) cr[k] = Array.isArray(v) ? Int32Array.from(v) : v;
    const g2 = await gw.create(cpu2, { rects: job.world.rects }, null, null, { creatures: cr });
    gw.upload(g2);
    const ncr = job.creatures.nCreatures;
    const ev = new Uint32Array(job.rolloutSteps * ncr);
    await gw.rollout(g2, job.law, 0, job.rolloutSteps, ev);
    await gw.download(g2);
    // per-creature readout after the rollout (6A+6 proprio inputs + 12 mantle floats)
    const stride = job.stateStride;
    const st = new Float32Array(ncr * stride);
    const wrote = gw.readState(g2, st, stride);
    out.rollout = Object.assign(state(cpu2), { events: Array.from(ev), state: Array.from(st), stateStride: wrote });
    gw.destroy(g2);
  }
  // respawn in place then one substep (train.js:188-229; gpuworld.js:344-372)
  if (job.respawn) {
    const cpu3 = makeWorld(job.world);
    const cr = {};
    for (const [k, v] of Object.entries(job.creatures)) cr[k] = Array.isArray(v) ? Int32Array.from(v) : v;
    const g3 = await gw.create(cpu3, { rects: job.world.rects }, null, null, { creatures: cr });
    gw.upload(g3);
    await gw.rollout(g3, job.law, 0, job.respawn.warm);
    gw.respawn(g3, job.respawn.creatures, job.respawn.poses);
    await gw.substep(g3);
    await gw.download(g3);
    out.respawn = state(cpu3);
    // handle lifetime: calls queue in order; a synchronous call on a busy
    // handle is refused; destroy while work is queued releases after it
    const k = require(path.join(__dirname, '..', 'node', 'kxpbd.node'));
    const h = g3.handle;
    const p1 = k.step(h, 2);
    const p2 = k.substep(h, 1);
    try { k.counts(h); out.busy = 'no throw'; } catch (e) { out.busy = String(e.message); }
    k.destroy(h);
    try { k.step(h, 1); out.dead = 'no throw'; } catch (e) { out.dead = String(e.message); }
    await p1;
    await p2;
    out.drained = true;
    g3.handle = null;
  }
  // error behaviour: a bad body index throws like the reference's guard
  const bad = makeWorld(job.world);
  bad.constraints[0].a = bad.n + 5;
  try {
    await gw.create(bad, { rects: job.world.rects });
    out.badThrows = false;
  } catch (e) {
    out.badThrows = true;
    out.badMessage = String(e.message);
  }
  return out;
}

// policy network and ES through node/brain.js: forward steps with the
// recurrent tail fed back, then ES generations with updates
function brainJob(job) {
  const B = require(path.join(__dirname, '..', 'node', 'brain.js'));
  const out = { nW: B.weightCount(job.nIn, job.hidden, B.OUT_DIM) };
  const pop = B.createPopulation(job.n, job.nIn, job.hidden, B.OUT_DIM);
  B.setGenomes(pop, Float32Array.from(job.genomes));
  const o = new Float32Array(job.n * B.OUT_DIM);
  const ctl = new Float32Array(job.n * 25);
  out.steps = [];
  for (const xs of job.inputs) {
    const x = Float32Array.from(xs);
    for (let i = 0; i < job.n; i++)
      for (let k = 0; k < B.OUT_DIM; k++) x[i * job.nIn + job.nIn - B.OUT_DIM + k] = o[i * B.OUT_DIM + k];
    B.forwardInto(pop, x, o, { controls: ctl, arms: 6 });
    out.steps.push({ out: Array.from(o), controls: Array.from(ctl) });
  }
  B.destroy(pop);
  const es = B.createES(out.nW, job.esOpts, Float32Array.from(job.theta));
  const epop = B.createPopulation(es.popSize, job.nIn, job.hidden, B.OUT_DIM);
  let st = B.trainerStreamState(job.seed);
  out.gens = [];
  for (const fit of job.fitness) {
    st = B.beginGeneration(es, epop, st);
    const genomes = Array.from(B.getGenomes(epop));
    const stats = B.update(es, fit);
    const s = B.esState(es);
    out.gens.push({ state: st, genomes, stats, theta: Array.from(s.theta), m: Array.from(s.m), v: Array.from(s.v),
      t: s.t, sigma: s.sigma });
  }
  // misuse throws like the rest of the surface
  try { B.beginGeneration(es, pop, 1); out.deadThrows = false; } catch (e) { out.deadThrows = String(e.message); }
  // a shape that is not the object's own is refused before any copy
  const small = Object.assign({}, epop, { n: 1 });
  try {
    B.forwardInto(small, new Float32Array(job.nIn), new Float32Array(B.OUT_DIM));
    out.shapeThrows = false;
  } catch (e) { out.shapeThrows = String(e.message); }
  try { B.update(Object.assign({}, es, { popSize: 2 }), [1, 2]); out.popThrows = false; } catch (e) { out.popThrows = String(e.message); }
  B.destroy(es);
  B.destroy(epop);
  return out;
}

(async () => {
  const [mode, file] = process.argv.slice(2);
  const job = JSON.parse(fs.readFileSync(file, 'utf8'), revive);
  const res = mode === 'pure' ? await pure(job) : mode === 'brain' ? brainJob(job) : await gpu(job);
  process.stdout.write(JSON.stringify(res) + '\n');
})().catch((e) => {
  process.stderr.write(String(e && e.stack ? e.stack : e) + '\n');
  process.exit(1);
});
