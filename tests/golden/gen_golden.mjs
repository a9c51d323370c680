// Golden-vector generator. Runs ONLY in the development container, never on a
// GPU host: it imports the reference implementation read-only from
// /root/reference (via esm_loader.mjs) and writes small JSON fixtures next to
// this file. The fixtures pin the oracle restatement under oracle/.
//
//   node --experimental-loader ./tests/golden/esm_loader.mjs tests/golden/gen_golden.mjs
//
// Node-12 safe: no top-level await, no optional chaining.
import { writeFileSync } from 'fs';
import { XpbdWorld, ensureColored, SLACK_POS } from '/root/reference/src/physics/xpbd.js';
import { makeMap, makeSdf, findSpawn, mulberry32 } from '/root/reference/src/world/map.js';
import { spawnKraken, attachGrip, tipPosition, adhesionToBreakForce, toneToCompliance,
  BODY, ARM_REACH } from '/root/reference/src/body/kraken.js';
import { createSim, applyControl, stepSim, readProprio } from '/root/reference/src/sim/core.js';
import { crnSpawns, respawnCreature } from '/root/reference/src/sim/train.js';
import { PROPRIO_DIM, SENSE_DIM } from '/root/reference/src/sim/core.js';
import { HIDDEN, OUT_DIM, inputDim, weightCount, randomGenome, forwardInto, decode } from '/root/reference/src/brain/net.js';
import { createES, genomeAt, update } from '/root/reference/src/brain/es.js';
import { f32ify } from '/root/reference/tools/f32solver.mjs';
import { packControls, mirrorInto, groupCollPoints, GPU_STRIDES } from '/root/reference/src/gpu/gpuworld.js';

const OUT = new URL('.', import.meta.url).pathname;
const dump = (name, obj) => {
  writeFileSync(OUT + name, JSON.stringify(obj));
  console.log('wrote', name);
};
const arr = (a, n) => Array.from(a.subarray ? a.subarray(0, n) : a.slice(0, n));

// ── control law shared with oracle/ and the device (counter-based mulberry32) ──
function mb32(a) {
  a = (a + 0x6D2B79F5) >>> 0;
  let t = a;
  t = Math.imul(t ^ (t >>> 15), t | 1);
  t ^= t + Math.imul(t ^ (t >>> 7), t | 61);
  return (t ^ (t >>> 14)) >>> 0;
}
function ctlU(seed, creature, step, slot) {
  let h = mb32(seed >>> 0);
  h = mb32((h ^ creature) >>> 0);
  h = mb32((h ^ step) >>> 0);
  h = mb32((h ^ slot) >>> 0);
  return h / 4294967296;
}
const SLOT_ADH = 256;
function lawControl(seed, ci, step, arms) {
  // per arm: shoulder target, per-joint targets, adhesion
  const out = [];
  for (let a = 0; a < arms.length; a++) {
    const nm = arms[a].motors.length;
    const tg = [];
    for (let m = 0; m < nm; m++) {
      const u = ctlU(seed, ci, step, a * 16 + m);
      const lim = m === 0 ? BODY.shoulderLimit : BODY.jointLimit;
      tg.push((2 * u - 1) * lim);
    }
    out.push({ targets: tg, adhesion: ctlU(seed, ci, step, SLOT_ADH + a) });
  }
  return out;
}
// creature record compatible with applyControl (core.js:183-220)
function reflexCreature(w, k) {
  return { world: w, arms: k.arms, perArm: k.arms.map(() => ({ shoulder: 0, curl: 0, adhesion: 0, cd: 0 })),
           opts: { gripSlipMax: 150, gripCooldown: 0.05 }, tone: 0.5 };
}
function applyLaw(cr, law) {
  applyControl(cr, { tone: 0.5, arms: law.map(l => ({ shoulder: l.targets[0], curl: 0, adhesion: l.adhesion })) });
  for (let a = 0; a < cr.arms.length; a++) {
    const ms = cr.arms[a].motors;
    for (let j = 1; j < ms.length; j++) {
      ms[j].target = Math.max(-BODY.jointLimit, Math.min(BODY.jointLimit, law[a].targets[j]));
    }
  }
}

// ── 1. PRNG streams ──
{
  const seeds = [0, 1, 7, 42, (1 ^ 0x9e3779b9) >>> 0, (71 ^ 0x5bf03635) >>> 0, 0xffffffff, 123456789];
  const streams = seeds.map(s => { const r = mulberry32(s); const v = []; for (let i = 0; i < 24; i++) v.push(r()); return { seed: s, values: v }; });
  const ctl = [];
  for (const s of [1, 7, 2024]) for (const c of [0, 1, 4095]) for (const st of [0, 1, 999]) for (const sl of [0, 5, 41, 256, 261]) {
    ctl.push([s, c, st, sl, ctlU(s, c, st, sl)]);
  }
  dump('prng.json', { streams, ctl });
}

// ── 2. maps ──
const mapCases = [];
for (const s of [1, 2, 3, 4, 5, 6]) mapCases.push([1200, 800, { seed: s }]);
for (const s of [3, 5, 7, 11, 13, 17, 23, 31, 47, 59, 71]) mapCases.push([1440, 810, { seed: s }]);
mapCases.push([4096, 4096, { corridorCols: 32, corridorRows: 32, seed: 1 }]);
mapCases.push([4096, 4096, { corridorCols: 32, corridorRows: 32, seed: 4 }]);
{
  const maps = mapCases.map(([w, h, o]) => {
    const m = makeMap(w, h, o);
    return { width: w, height: h, opts: o, rects: m.rects.map(r => [r.x, r.y, r.w, r.h]) };
  });
  dump('maps.json', { maps });
}

// ── 3. SDF queries incl. the best² quirk ──
{
  const map = makeMap(1200, 800, { seed: 1 });
  const rnd = mulberry32(555);
  const pts = [];
  for (let i = 0; i < 600; i++) pts.push([-30 + rnd() * 1260, -30 + rnd() * 860]);
  for (const r of map.rects) {      // inside, on edges, at corners, just outside
    const cx = r.x + r.w / 2, cy = r.y + r.h / 2;
    pts.push([cx, cy], [r.x, cy], [r.x + r.w, cy], [cx, r.y], [cx, r.y + r.h], [r.x, r.y], [r.x + r.w, r.y + r.h],
      [r.x - 3.5, cy], [cx, r.y + r.h + 2.25], [r.x + 1e-7, r.y + 1e-7], [r.x + r.w * 0.25, r.y + r.h * 0.5]);
  }
  const q = (sdf, p) => { const h = sdf(p[0], p[1]); return h ? [p[0], p[1], h.d, h.nx, h.ny] : [p[0], p[1], null]; };
  const main = pts.map(p => q(map.sdf, p));
  // quirk: an interior hit (best<0) is overridden by a later exterior rect
  // closer than |best| because the exterior test compares d2 with best².
  const quirkRects = [{ x: 0, y: 0, w: 40, h: 40 }, { x: 45, y: 10, w: 10, h: 10 },
                      { x: 100, y: 100, w: 20, h: 5 }, { x: 30, y: 50, w: 5, h: 5 }];
  const qs = makeSdf(quirkRects);
  const qpts = [[35, 15], [38, 15], [20, 20], [33, 45], [33, 38], [50, 25], [60, 60], [110, 102], [42.5, 15], [39.9, 39.9]];
  const quirk = qpts.map(p => q(qs, p));
  const empty = q(makeSdf([]), [1, 2]);
  dump('sdf.json', { map: { width: 1200, height: 800, seed: 1 }, main, quirk: { rects: quirkRects.map(r => [r.x, r.y, r.w, r.h]), queries: quirk }, empty });
}

// ── 4. findSpawn sequence (createSim, core.js:42-50) ──
{
  const out = [];
  for (const [w, h, seed, n] of [[1200, 800, 1, 64], [1440, 810, 7, 8], [4096, 4096, 1, 16]]) {
    const map = w === 4096 ? makeMap(w, h, { corridorCols: 32, corridorRows: 32, seed }) : makeMap(w, h, { seed });
    const rnd = mulberry32(seed ^ 0x9e3779b9);
    const sp = [];
    for (let i = 0; i < n; i++) { const p = findSpawn(map, ARM_REACH + 20, rnd); sp.push([p.x, p.y, rnd() * Math.PI * 2]); }
    out.push({ width: w, height: h, seed, spawns: sp });
  }
  dump('spawns.json', { cases: out });
}

// ── 5. spawnKraken + ensureColored ──
function worldDump(w) {
  return {
    n: w.n, px: arr(w.px, w.n), py: arr(w.py, w.n), ang: arr(w.ang, w.n), im: arr(w.im, w.n), ii: arr(w.ii, w.n),
    cons: w.constraints.map(c => [c.type, c.a, c.b, c.lax, c.lay, c.lbx, c.lby, c.compliance, c.rest,
      Number.isFinite(c.minAngle) ? c.minAngle : null, Number.isFinite(c.maxAngle) ? c.maxAngle : null,
      c.owner, c.broken ? 1 : 0, Number.isFinite(c.breakForce) ? c.breakForce : null, c.target, c.color]),
    pts: w.collPts.map(p => [p.body, p.lx, p.ly, p.r]),
  };
}
{
  const cases = [];
  for (const [x, y, a, nk] of [[600, 400, 0.7, 1], [300.5, 200.25, 5.9, 1], [720, 405, 0.0, 3]]) {
    const w = new XpbdWorld({ substeps: 4, iterations: 8, h: 1 / 240 });
    for (let i = 0; i < nk; i++) spawnKraken(w, x + 250 * i, y - 40 * i, a + 0.3 * i);
    const creation = worldDump(w);
    const ids = new Map(); w.constraints.forEach((c, i) => ids.set(c, i));
    ensureColored(w);
    cases.push({ spawn: [x, y, a, nk], creation, colors: creation.cons.map((_, i) => null),
      sorted: w.constraints.map(c => ids.get(c)), colorOf: w.constraints.map(c => c.color),
      colorOff: arr(w.colorOff, 16), colorCnt: arr(w.colorCnt, 16) });
  }
  dump('kraken.json', { body: BODY, armReach: ARM_REACH, cases,
    toneToCompliance: [0, 0.25, 0.5, 1, -1, 2].map(t => [t, toneToCompliance(t)]),
    adhesionToBreakForce: [0, 0.05, 0.5, 1, -1, 2].map(t => [t, adhesionToBreakForce(t)]) });
}

// ── 6. trajectories under the seeded control law: f64 oracle and f32 emulator ──
function buildTrajWorld(S, I, h) {
  const map = makeMap(1200, 800, { seed: 1 });
  const w = new XpbdWorld({ substeps: S, iterations: I, h });
  w.sdf = map.sdf;
  const ks = [];
  // one creature hugging the left border wall, one beside the first corridor
  // wall, one in the open — collisions, grips and limits all get exercised
  const r4 = map.rects[4];
  const poses = [[60, 300, 0.3], [r4.x + r4.w + 30, r4.y + Math.min(r4.h, 200) * 0.5, 2.1], [600, 420, 4.4]];
  for (const p of poses) ks.push(spawnKraken(w, p[0], p[1], p[2]));
  ensureColored(w);
  return { w, map, ks, poses };
}
function stateDump(w, ks, crs) {
  return {
    px: arr(w.px, w.n), py: arr(w.py, w.n), ang: arr(w.ang, w.n), vx: arr(w.vx, w.n), vy: arr(w.vy, w.n), av: arr(w.av, w.n),
    grips: ks.map(k => k.arms.map(a => [a.grip.broken ? 1 : 0, a.grip.lbx, a.grip.lby,
      Number.isFinite(a.grip.breakForce) ? a.grip.breakForce : null, a.grip.lambdaX, a.grip.lambdaY, a.grip.force])),
    cd: crs.map(c => c.perArm.map(p => p.cd)),
    lambda: w.constraints.map(c => c.lambda),
  };
}
function runTraj(mode, S, I, h, nSteps, record, seed) {
  const { w, map, ks, poses } = buildTrajWorld(S, I, h);
  let emu = null;
  if (mode === 'f32') { emu = f32ify(w, map); emu.roundState(); }
  const crs = ks.map(k => reflexCreature(w, k));
  const frames = [];
  const events = [];
  for (let s = 0; s < nSteps; s++) {
    const before = ks.map(k => k.arms.map(a => a.grip.broken));
    for (let ci = 0; ci < ks.length; ci++) applyLaw(crs[ci], lawControl(seed, ci, s, ks[ci].arms));
    const mid = ks.map(k => k.arms.map(a => a.grip.broken));
    w.step();
    for (let ci = 0; ci < ks.length; ci++) for (let a = 0; a < 6; a++) {
      if (before[ci][a] && !mid[ci][a]) events.push([s, ci, a, 'attach']);
      if (!before[ci][a] && mid[ci][a]) events.push([s, ci, a, 'release']);
      if (!mid[ci][a] && ks[ci].arms[a].grip.broken) events.push([s, ci, a, 'break']);
    }
    if (record.indexOf(s + 1) >= 0) frames.push({ step: s + 1, state: stateDump(w, ks, crs) });
  }
  return { mode, S, I, h, seed, poses, frames, events };
}
{
  const rec = [1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64];
  const t64 = runTraj('f64', 4, 8, 1 / 240, 64, rec, 2024);
  const t32 = runTraj('f32', 4, 8, 1 / 240, 64, rec, 2024);
  const tdef = runTraj('f64', 2, 4, 1 / 120, 32, [1, 2, 4, 8, 16, 32], 77);
  dump('traj.json', { traj: [t64, t32, tdef] });
}

// ── 7. one substep, phase by phase (integrate → constraints → collisions) ──
{
  const { w, ks } = buildTrajWorld(4, 8, 1 / 240);
  const crs = ks.map(k => reflexCreature(w, k));
  for (let ci = 0; ci < ks.length; ci++) applyLaw(crs[ci], lawControl(99, ci, 0, ks[ci].arms));
  // give every body a velocity so the prediction is non-trivial
  const rnd = mulberry32(31337);
  for (let i = 0; i < w.n; i++) { w.vx[i] = (rnd() - 0.5) * 200; w.vy[i] = (rnd() - 0.5) * 200; w.av[i] = (rnd() - 0.5) * 10; }
  const init = stateDump(w, ks, crs);
  w.substep();
  dump('substep.json', { seed: 99, init, after: stateDump(w, ks, crs), lambdaX: w.constraints.map(c => c.lambdaX), lambdaY: w.constraints.map(c => c.lambdaY) });
}

// ── 8. harness KAT scenarios (harness.js:46-251), reference defaults S=2,I=4,h=1/120 ──
{
  const kat = {};
  {
    const sim = createSim({ seed: 7, creatures: 1 });
    const c = sim.creatures[0];
    for (let i = 0; i < 960; i++) stepSim(sim);
    kat.passive = { pathLen: c.pathLen, disp: Math.hypot(c.x - c.spawn.x, c.y - c.spawn.y), maxAngSpeed: sim.world.stats.maxAngSpeed, nan: sim.world.stats.nan,
      spawn: [c.spawn.x, c.spawn.y, c.spawn.angle] };
  }
  {
    const rows = [];
    for (const tone of [0.3, 0.65, 1.0]) {
      const sim = createSim({ seed: 11, creatures: 1 });
      const c = sim.creatures[0], arm = c.arms[0];
      attachGrip(sim.world, arm, adhesionToBreakForce(1.0));
      const anchor = Object.assign({}, tipPosition(sim.world, arm));
      const startDist = Math.hypot(sim.world.px[c.mantle] - anchor.x, sim.world.py[c.mantle] - anchor.y);
      let minDist = startDist, broke = false, steps = 0;
      for (let i = 0; i < 600; i++) {
        applyControl(c, { tone, arms: c.arms.map((_, j) => j === 0 ? { shoulder: 0, curl: BODY.jointLimit * 0.98, adhesion: 1 } : { shoulder: 0, curl: 0, adhesion: 0 }) });
        stepSim(sim); steps++;
        if (arm.grip.broken) { broke = true; break; }
        const d = Math.hypot(sim.world.px[c.mantle] - anchor.x, sim.world.py[c.mantle] - anchor.y);
        if (d < minDist) minDist = d;
      }
      rows.push({ tone, startDist, pulled: startDist - minDist, broke, steps, final: [sim.world.px[c.mantle], sim.world.py[c.mantle]] });
    }
    kat.grip = rows;
  }
  {
    const res = {};
    for (const [tag, adh] of [['weak', 0.05], ['strong', 1.0]]) {
      const sim = createSim({ seed: 13, creatures: 1 });
      const c = sim.creatures[0], arm = c.arms[0];
      attachGrip(sim.world, arm, adhesionToBreakForce(adh));
      let broke = false, steps = 0;
      for (let i = 0; i < 400; i++) {
        applyControl(c, { tone: 1.0, arms: c.arms.map((_, j) => j === 0 ? { shoulder: 0, curl: BODY.jointLimit * 0.9, adhesion: adh }
          : { shoulder: (j % 2 ? 1 : -1) * BODY.shoulderLimit * 0.8, curl: BODY.jointLimit * 0.9, adhesion: 0 }) });
        stepSim(sim); steps++;
        if (arm.grip.broken) { broke = true; break; }
      }
      res[tag] = { broke, steps, force: arm.grip.force || 0, limit: adhesionToBreakForce(adh) };
    }
    kat.break = res;
  }
  {
    const map = makeMap(1440, 810, { seed: 5 });
    const sp = findSpawn(map, ARM_REACH + 20, mulberry32(9));
    const w1 = new XpbdWorld(); w1.sdf = map.sdf;
    const k1 = spawnKraken(w1, sp.x, sp.y, 0.4);
    w1.px[k1.arms[0].segs[1]] += SLACK_POS * 0.4;
    const px0 = Float64Array.from(w1.px);
    for (let i = 0; i < 240; i++) w1.step();
    let moved = 0; for (let i = 0; i < w1.n; i++) moved = Math.max(moved, Math.abs(w1.px[i] - px0[i]));
    const w2 = new XpbdWorld(); w2.sdf = map.sdf;
    const k2 = spawnKraken(w2, sp.x, sp.y, 0.4);
    w2.px[k2.arms[0].segs[1]] += 0.05;
    for (let i = 0; i < 240; i++) w2.step();
    kat.slack = { spawn: [sp.x, sp.y], fixpointMoved: moved, fixpointMaxSpeed: w1.stats.maxSpeed,
      corrected: { px: arr(w2.px, w2.n), py: arr(w2.py, w2.n), ang: arr(w2.ang, w2.n) } };
  }
  dump('kat.json', kat);
}

// ── 9. gpuworld pure helpers: packControls / mirrorInto / groupCollPoints ──
{
  const sim = createSim({ seed: 33, creatures: 2 });
  const w = sim.world, nC = w.constraints.length, nB = w.n;
  for (let k = 0; k < nC; k++) {
    const c = w.constraints[k];
    c.compliance = 0.123 + k * 1e-4; c.target = 0.01 * k;
    c.breakForce = k % 3 === 0 ? Infinity : (k % 3 === 1 ? 5 + k : 3e38);
    c.broken = k % 2 === 0; c.lbx = 11 + k * 1.1; c.lby = -22 - k * 0.7;
  }
  const cd = new Float32Array(nC * GPU_STRIDES.constraint);
  packControls(w, cd);
  const packed = Array.from(new Uint32Array(cd.buffer));
  const bd = new Float32Array(nB * GPU_STRIDES.body);
  const cdm = new Float32Array(nC * GPU_STRIDES.constraint);
  const cu = new Uint32Array(cdm.buffer);
  for (let i = 0; i < nB; i++) {
    const o = i * 12;
    bd[o] = 100 + i * 0.5; bd[o + 1] = -200 + i * 0.25; bd[o + 4] = 1 + i * 0.01; bd[o + 5] = -2 - i * 0.01;
    bd[o + 6] = 0.3 + i * 0.001; bd[o + 8] = 0.05 - i * 0.002;
  }
  bd[12 * 5 + 1] = NaN;
  for (let k = 0; k < nC; k++) { const o = k * 20; cdm[o + 13] = 0.5 + k * 0.001; cdm[o + 14] = -0.25 + k * 0.001; cdm[o + 15] = k * 0.002; cu[o + 17] = k % 2; }
  const stats = mirrorInto(w, bd, cdm);
  const mirrored = { px: arr(w.px, nB).map(v => (Number.isNaN(v) ? 'NaN' : v)), py: arr(w.py, nB).map(v => (Number.isNaN(v) ? 'NaN' : v)),
    lambda: w.constraints.map(c => c.lambda), broken: w.constraints.map(c => c.broken ? 1 : 0),
    force: w.constraints.map(c => c.force), stats };
  const types = w.constraints.map(c => c.type);
  const pts = w.collPts.map(p => p.body);
  const rnd = mulberry32(77);
  const shuffled = w.collPts.map((p, i) => ({ body: p.body, lx: p.lx, ly: p.ly, r: p.r, i }));
  for (let i = shuffled.length - 1; i > 0; i--) { const j = (rnd() * (i + 1)) | 0; const t = shuffled[i]; shuffled[i] = shuffled[j]; shuffled[j] = t; }
  const g1 = groupCollPoints(w.collPts, nB), g2 = groupCollPoints(shuffled, nB);
  dump('gpuworld.json', { strides: GPU_STRIDES, nB, nC, h: w.h, types, packed,
    bodyIn: Array.from(new Uint32Array(bd.buffer)), consIn: Array.from(new Uint32Array(cdm.buffer)), mirrored,
    group: { bodies: pts, order: Array.from(g1.order), range: Array.from(g1.range),
      shuffledBodies: shuffled.map(p => p.body), shuffledOrder: Array.from(g2.order), shuffledRange: Array.from(g2.range) } });
}
// ── 10. proprioception (core.js:129-175 readProprio) after every step of the
// trajectory world under the seeded law, the acceleration input following the
// previous call's mantle velocity as in the sim loop ──
{
  const S = 4, I = 8, h = 1 / 240, seed = 2024, nSteps = 32, rec = [1, 2, 4, 8, 16, 32];
  const { w, ks } = buildTrajWorld(S, I, h);
  const crs = ks.map(k => reflexCreature(w, k));
  const cs = ks.map(k => ({ world: w, mantle: k.mantle, arms: k.arms, prevVx: 0, prevVy: 0 }));
  const frames = [];
  for (let s = 0; s < nSteps; s++) {
    for (let ci = 0; ci < ks.length; ci++) applyLaw(crs[ci], lawControl(seed, ci, s, ks[ci].arms));
    w.step();
    const prev = cs.map(c => [c.prevVx, c.prevVy]);
    const outs = cs.map(c => Array.from(readProprio(c, new Float64Array(42))));
    if (rec.indexOf(s + 1) >= 0) {
      frames.push({ step: s + 1, px: arr(w.px, w.n), py: arr(w.py, w.n), ang: arr(w.ang, w.n),
        vx: arr(w.vx, w.n), vy: arr(w.vy, w.n), av: arr(w.av, w.n),
        grips: ks.map(k => k.arms.map(a => [a.grip.broken ? 1 : 0, a.grip.force,
          Number.isFinite(a.grip.breakForce) ? a.grip.breakForce : null])),
        prev, proprio: outs,
        side: cs.map(c => [c.meanExtension, c.meanCurl, c.anchoredCount, c.speed]) });
    }
  }
  dump('proprio.json', { S, I, h, seed, armReach: ARM_REACH, segLen: BODY.segLen,
    shoulderLimit: BODY.shoulderLimit, jointLimit: BODY.jointLimit,
    creatures: ks.map(k => ({ mantle: k.mantle, arms: k.arms.map(a => a.segs) })), frames });
}
// ── 11. policy network (net.js:71-124 forwardInto, decode) over a small
// population with per-creature genomes and the recurrent output feedback of
// gputrain.js:77-86, then the ES generation (es.js:67-149 genomeAt, update)
// driven by the trainer's stream (train.js:65-72 is mulberry32 seeded with
// seed*2654435761) ──
{
  const b64 = (a) => Buffer.from(a.buffer, a.byteOffset, a.byteLength).toString('base64');
  const nIn = inputDim(PROPRIO_DIM, SENSE_DIM), nW = weightCount(nIn);
  const pop = 6, steps = 3, arms = 6;
  const rnd = mulberry32(12345);
  const genomes = new Float32Array(pop * nW);
  for (let i = 0; i < pop; i++) {
    const g = randomGenome(nIn, HIDDEN, OUT_DIM, rnd);
    const sc = i >= 4 ? 25 : 1;                      // two genomes large enough to hit the +-20 clamp
    for (let k = 0; k < nW; k++) genomes[i * nW + k] = g[k] * sc;
  }
  const outputs = new Float32Array(pop * OUT_DIM);
  const scratch = new Float32Array(pop * 2 * HIDDEN);
  const frames = [];
  for (let s = 0; s < steps; s++) {
    const inputs = new Float32Array(pop * nIn);
    for (let i = 0; i < pop; i++) {
      const x = inputs.subarray(i * nIn, (i + 1) * nIn);
      for (let k = 0; k < PROPRIO_DIM + SENSE_DIM; k++) x[k] = (rnd() * 2 - 1) * (k % 7 === 0 ? 4 : 1.5);
      const oPrev = outputs.subarray(i * OUT_DIM, (i + 1) * OUT_DIM);
      for (let k = 0; k < OUT_DIM; k++) x[PROPRIO_DIM + SENSE_DIM + k] = oPrev[k];
      forwardInto(genomes.subarray(i * nW, (i + 1) * nW), x, oPrev,
        scratch.subarray(i * 2 * HIDDEN, (i + 1) * 2 * HIDDEN), nIn, HIDDEN, OUT_DIM);
    }
    const controls = [];
    for (let i = 0; i < pop; i++) {
      const c = { tone: 0, emitEcho: false, arms: [] };
      for (let a = 0; a < arms; a++) c.arms.push({ shoulder: 0, curl: 0, adhesion: 0 });
      decode(outputs.subarray(i * OUT_DIM, (i + 1) * OUT_DIM), c, arms);
      controls.push([c.tone].concat(...c.arms.map(a => [a.shoulder, a.curl, a.adhesion]), [c.emitEcho ? 1 : 0]));
    }
    frames.push({ inputs: b64(inputs), outputs: b64(outputs), scratch: b64(scratch), controls });
  }
  dump('net.json', { nIn, hidden: HIDDEN, nOut: OUT_DIM, nW, pop, proprioDim: PROPRIO_DIM, senseDim: SENSE_DIM,
    genomes: b64(genomes), frames });

  const popSize = 4, seed = 7;
  const es = createES(nW, { popSize });
  es.theta.set(randomGenome(nIn, HIDDEN, OUT_DIM, mulberry32(99)));
  const theta0 = Float32Array.from(es.theta);
  const s0 = (seed * 2654435761) >>> 0;
  const trnd = mulberry32(s0);
  const gens = [];
  const frnd = mulberry32(31337);
  for (let gen = 0; gen < 2; gen++) {
    const sigma = es.opts.sigma;
    const gs = new Float32Array(popSize * nW);
    for (let i = 0; i < popSize; i++) genomeAt(es, i, gs.subarray(i * nW, (i + 1) * nW), trnd);
    const fitness = new Float64Array(popSize);
    for (let i = 0; i < popSize; i++) fitness[i] = (frnd() - 0.3) * 40;
    const stats = update(es, fitness);
    gens.push({ sigma, genomes: b64(gs), fitness: Array.from(fitness), theta: b64(es.theta), m: b64(es.m),
      v: b64(es.v), t: es.t, stats });
  }
  dump('es.json', { nW, popSize, seed, s0, opts: es.opts, theta0: b64(theta0), gens });
}
// ── 12. generalised morphology. spawnKraken reads BODY.arms and
// BODY.segsPerArm at call time (kraken.js:64-148; the mantle keeps its 8
// collision points), so setting them gives the reference's own creature for
// each (arms, segs): bodies, points, constraints, the greedy colouring, a
// trajectory under the seeded law with its grip events (f64, and the f32
// emulator for two shapes) and readProprio (core.js:129-175) after each
// recorded step. BODY is restored afterwards. ──
{
  const saved = [BODY.arms, BODY.segsPerArm];
  const morphs = [[4, 5], [4, 9], [5, 8], [7, 6], [8, 9], [8, 5], [5, 5], [6, 7]];
  const cases = [];
  const S = 4, I = 8, h = 1 / 240, seed = 4242, nSteps = 24, rec = [1, 8, 24];
  for (const [A, SG] of morphs) {
    BODY.arms = A; BODY.segsPerArm = SG;
    for (const mode of (A === 8 && SG === 9) || (A === 4 && SG === 5) ? ['f64', 'f32'] : ['f64']) {
      const map = makeMap(1200, 800, { seed: 1 });
      const w = new XpbdWorld({ substeps: S, iterations: I, h });
      w.sdf = map.sdf;
      const r4 = map.rects[4];
      const poses = [[70, 300, 0.3], [r4.x + r4.w + 40, r4.y + Math.min(r4.h, 200) * 0.5, 2.1], [600, 420, 4.4]];
      const ks = poses.map(p => spawnKraken(w, p[0], p[1], p[2]));
      const creation = worldDump(w);
      const ids = new Map(); w.constraints.forEach((c, i) => ids.set(c, i));
      ensureColored(w);
      const colouring = { sorted: w.constraints.map(c => ids.get(c)), colorOf: w.constraints.map(c => c.color),
        colorOff: arr(w.colorOff, 16), colorCnt: arr(w.colorCnt, 16) };
      let emu = null;
      if (mode === 'f32') { emu = f32ify(w, map); emu.roundState(); }
      const crs = ks.map(k => reflexCreature(w, k));
      const cs = ks.map(k => ({ world: w, mantle: k.mantle, arms: k.arms, prevVx: 0, prevVy: 0 }));
      const frames = [], events = [];
      for (let s = 0; s < nSteps; s++) {
        const before = ks.map(k => k.arms.map(a => a.grip.broken));
        for (let ci = 0; ci < ks.length; ci++) applyLaw(crs[ci], lawControl(seed, ci, s, ks[ci].arms));
        const mid = ks.map(k => k.arms.map(a => a.grip.broken));
        w.step();
        for (let ci = 0; ci < ks.length; ci++) for (let a = 0; a < A; a++) {
          if (before[ci][a] && !mid[ci][a]) events.push([s, ci, a, 'attach']);
          if (!before[ci][a] && mid[ci][a]) events.push([s, ci, a, 'release']);
          if (!mid[ci][a] && ks[ci].arms[a].grip.broken) events.push([s, ci, a, 'break']);
        }
        const prev = cs.map(c => [c.prevVx, c.prevVy]);
        const outs = mode === 'f64' ? cs.map(c => Array.from(readProprio(c, new Float64Array(6 * A + 6)))) : null;
        if (rec.indexOf(s + 1) >= 0) {
          const fr = { step: s + 1, state: stateDump(w, ks, crs) };
          if (outs) { fr.prev = prev; fr.proprio = outs; fr.side = cs.map(c => [c.meanExtension, c.meanCurl, c.anchoredCount, c.speed]); }
          frames.push(fr);
        }
      }
      cases.push({ arms: A, segs: SG, mode, S, I, h, seed, poses, mapSeed: 1, creation: mode === 'f64' ? creation : null,
        colouring: mode === 'f64' ? colouring : null, frames, events,
        creatures: ks.map(k => ({ mantle: k.mantle, arms: k.arms.map(a => a.segs) })) });
    }
  }
  BODY.arms = saved[0]; BODY.segsPerArm = saved[1];
  dump('morph.json', { armReach: ARM_REACH, cases });
}

// ── 13. crnSpawns (train.js:103-122) drawn from the trainer's stream
// (train.js:65-72: mulberry32 seeded with seed*2654435761), pair-shared and
// not, and respawnCreature (train.js:188-229) on a running world: the full
// body and constraint state before and after the respawn, then 8 further steps
// under the seeded law ──
{
  const crn = [];
  for (const [W, H, o, seed, pop, pairShared] of [[1200, 800, { seed: 1 }, 1, 64, true], [1440, 810, { seed: 7 }, 3, 17, false],
    [4096, 4096, { corridorCols: 32, corridorRows: 32, seed: 1 }, 2, 48, true], [1200, 800, { seed: 2 }, 9, 40, true]]) {
    const map = makeMap(W, H, o);
    const s0 = (seed * 2654435761) >>> 0;
    const rnd = mulberry32(s0);
    let calls = 0;
    const counted = () => { calls++; return rnd(); };
    const sp = crnSpawns({ map }, pop, counted, pairShared);
    crn.push({ width: W, height: H, opts: o, seed, s0, pop, pairShared, calls, spawns: sp.map(p => [p.x, p.y, p.angle]) });
  }
  {
    const rnd = mulberry32(5);
    const sp = crnSpawns({ map: null }, 6, rnd, true);
    crn.push({ width: 240, height: 240, opts: null, seed: null, s0: 5, pop: 6, pairShared: true, calls: null, spawns: sp.map(p => [p.x, p.y, p.angle]) });
  }
  const full = (w) => ({
    px: arr(w.px, w.n), py: arr(w.py, w.n), ppx: arr(w.ppx, w.n), ppy: arr(w.ppy, w.n), vx: arr(w.vx, w.n), vy: arr(w.vy, w.n),
    ang: arr(w.ang, w.n), angP: arr(w.angP, w.n), av: arr(w.av, w.n),
    cons: w.constraints.map(c => [c.target, c.broken ? 1 : 0, c.force || 0, c.lambdaX, c.lambdaY, c.lambda, c.compliance,
      c.lbx, c.lby, Number.isFinite(c.breakForce) ? c.breakForce : null]),
  });
  const resp = [];
  const saved = [BODY.arms, BODY.segsPerArm];
  for (const [A, SG, seed] of [[6, 7, 21], [5, 8, 22]]) {
    BODY.arms = A; BODY.segsPerArm = SG;
    const sim = createSim({ seed, creatures: 3, substeps: 4, iterations: 8, h: 1 / 240, width: 1200, height: 800 });
    const w = sim.world;
    const spawns = sim.creatures.map(c => [c.spawn.x, c.spawn.y, c.spawn.angle]);
    const law = 777;
    const run = (s0, n) => {
      for (let s = s0; s < s0 + n; s++) {
        for (let ci = 0; ci < sim.creatures.length; ci++) applyLaw(sim.creatures[ci], lawControl(law, ci, s, sim.creatures[ci].arms));
        w.step();
      }
    };
    run(0, 40);
    const before = full(w);
    const cdBefore = sim.creatures.map(c => c.perArm.map(p => p.cd));
    const map = sim.map;
    const targets = crnSpawns(sim, 4, mulberry32(99), true);
    respawnCreature(sim, sim.creatures[0], targets[0].x, targets[0].y, targets[0].angle);
    respawnCreature(sim, sim.creatures[2], targets[2].x, targets[2].y, targets[2].angle);
    const after = full(w);
    const cdAfter = sim.creatures.map(c => c.perArm.map(p => p.cd));
    const prevAfter = sim.creatures.map(c => [c.prevVx, c.prevVy]);
    run(40, 8);
    resp.push({ arms: A, segs: SG, seed, law, spawns, mapRects: map.rects.map(r => [r.x, r.y, r.w, r.h]),
      respawned: [[0, targets[0].x, targets[0].y, targets[0].angle], [2, targets[2].x, targets[2].y, targets[2].angle]],
      before, after, cdBefore, cdAfter, prevAfter, later: full(w), laterCd: sim.creatures.map(c => c.perArm.map(p => p.cd)),
      colorOf: w.constraints.map(c => c.color) });
  }
  BODY.arms = saved[0]; BODY.segsPerArm = saved[1];
  dump('respawn.json', { spawnClear: 26, armReach: ARM_REACH, crn, respawn: resp });
}
console.log('done');
