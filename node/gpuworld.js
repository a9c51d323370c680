'use strict';
// The gpuworld surface (src/gpu/gpuworld.js of the reference) over libkxpbd on
// an MI355X: same exported names, argument meaning, side effects on the
// caller's XpbdWorld and error behaviour, so a Backend built on it
// ({upload, uploadControls, step, download}, src/sim/gputrain.js:31-47) runs
// unchanged. The device work is HIP (libkxpbd.so) through kxpbd.node.
//
//   create(cpu, map, wgslSource?, device?, opts?) -> Promise<g>
//   upload(g), uploadControls(g), packControls(cpu, consData)
//   step(g) -> Promise, substep(g) -> Promise,
//   download(g) -> Promise<{maxSpeed, maxAngSpeed, nan}>
//   respawn(g, creatures, poses), rollout(g, law, step0, n, events?), readState
//   mirrorInto(cpu, bodyData, consData) -> stats, groupCollPoints(collPts, n)
//   destroy(g), GPU_STRIDES, isAvailable()
//
// wgslSource and device are accepted and ignored (the kernels are compiled
// HIP); opts = {path: 'persistent'|'plain', lanesPerCreature, wavesPerBlock}.
const path = require('path');

const GPU_STRIDES = Object.freeze({ body: 12, constraint: 20, collPt: 4, rect: 4, force: 4 });
const C_PIN = 0;
const C_ANGLE = 1;
const C_GRIP = 2;
const MAX_COLORS = 16;
const PATHS = { plain: 0, persistent: 1 };

let native = null;
function lib() {
  if (!native) native = require(path.join(__dirname, 'kxpbd.node'));
  return native;
}

function isAvailable() {
  try {
    lib();
    return true;
  } catch (e) {
    return false;
  }
}

function f32(n, fill) {
  const a = new Float32Array(n);
  if (fill !== undefined) a.fill(fill);
  return a;
}

// breakForce as the device stores it: +inf -> 1e30 (gpuworld.js:336)
function packedBreak(v) {
  return Number.isFinite(v) ? Math.min(v, 1e30) : 1e30;
}

/** gpuworld.create: colours the world (ensureColored semantics: c.color set,
 * constraints stably grouped by colour, colorOff/colorCnt/colorMax stored),
 * builds the device world. Throws Error on implausible parameters or bad
 * indices, as the reference's sanity guards do. */
async function create(cpu, map, wgslSource, device, opts) {
  const o = opts || {};
  const k = lib();
  const n = cpu.n;
  const cons = cpu.constraints;
  const nc = cons.length;
  const pts = cpu.collPts;
  const np = pts.length;
  const rects = (map && map.rects) || [];
  const nr = rects.length;
  const desc = {
    nBodies: n, nConstraints: nc, nPoints: np, nRects: nr,
    invMass: f32(n), invInertia: f32(n),
    cType: new Int32Array(nc), cA: new Int32Array(nc), cB: new Int32Array(nc), cOwner: new Int32Array(nc),
    cLax: f32(nc), cLay: f32(nc), cLbx: f32(nc), cLby: f32(nc),
    cRest: f32(nc), cMin: f32(nc), cMax: f32(nc), cColor: new Int32Array(nc),
    pBody: new Int32Array(np), pLx: f32(np), pLy: f32(np), pR: f32(np),
    rects: f32(4 * nr),
  };
  for (let i = 0; i < n; i++) {
    desc.invMass[i] = cpu.im[i];
    desc.invInertia[i] = cpu.ii[i];
  }
  for (let i = 0; i < nc; i++) {
    const c = cons[i];
    desc.cType[i] = c.type;
    desc.cA[i] = c.a;
    desc.cB[i] = c.b >= 0 ? c.b : -1;
    desc.cOwner[i] = c.owner !== undefined ? c.owner : -1;
    desc.cLax[i] = c.lax; desc.cLay[i] = c.lay;
    desc.cLbx[i] = c.lbx; desc.cLby[i] = c.lby;
    desc.cRest[i] = c.rest || 0;
    desc.cMin[i] = c.minAngle !== undefined ? c.minAngle : -Infinity;
    desc.cMax[i] = c.maxAngle !== undefined ? c.maxAngle : Infinity;
    desc.cColor[i] = c.color !== undefined ? c.color : -1;
  }
  for (let i = 0; i < np; i++) {
    const p = pts[i];
    desc.pBody[i] = p.body; desc.pLx[i] = p.lx; desc.pLy[i] = p.ly; desc.pR[i] = p.r;
  }
  for (let i = 0; i < nr; i++) {
    const r = rects[i];
    desc.rects[4 * i] = r.x; desc.rects[4 * i + 1] = r.y;
    desc.rects[4 * i + 2] = r.w; desc.rects[4 * i + 3] = r.h;
  }
  const params = {
    substeps: cpu.substeps, iterations: cpu.iterations, h: cpu.h,
    linDrag: cpu.linDrag, angDrag: cpu.angDrag, maxSpeed: cpu.maxSpeed, maxAngSpeed: cpu.maxAngSpeed,
    path: PATHS[o.path || 'persistent'],
    lanesPerCreature: o.lanesPerCreature || 0, wavesPerBlock: o.wavesPerBlock || 0,
  };
  const handle = k.create(desc, params, o.creatures);
  // ensureColored (xpbd.js:108-126) on the caller's world
  const order = new Int32Array(nc);
  const color = new Int32Array(nc);
  k.constraintOrder(handle, { order, color });
  const sorted = new Array(nc);
  for (let i = 0; i < nc; i++) {
    const c = cons[order[i]];
    c.color = color[i];
    sorted[i] = c;
  }
  cpu.constraints = sorted;
  const off = new Uint32Array(MAX_COLORS);
  const cnt = new Uint32Array(MAX_COLORS);
  k.colorInfo(handle, { off, cnt });
  cpu.colorOff = off;
  cpu.colorCnt = cnt;
  cpu.colorMax = Math.max(...cnt);
  const g = {
    cpu, map, handle, nB: n, nC: nc, nP: np, nR: nr,
    dispatches: 0,
    consData: new Float32Array(nc * GPU_STRIDES.constraint),
    body: { px: f32(n), py: f32(n), vx: f32(n), vy: f32(n), ang: f32(n), av: f32(n),
            fx: f32(n), fy: f32(n), tq: f32(n) },
    cons: { lbx: f32(nc), lby: f32(nc), compliance: f32(nc), target: f32(nc), breakForce: f32(nc),
            broken: new Uint8Array(nc), lambda: f32(nc), lambdaX: f32(nc), lambdaY: f32(nc) },
  };
  return g;
}

function fillControls(g) {
  const cs = g.cpu.constraints;
  const c = g.cons;
  for (let i = 0; i < g.nC; i++) {
    const k = cs[i];
    c.lbx[i] = k.lbx; c.lby[i] = k.lby;
    c.compliance[i] = k.compliance; c.target[i] = k.target || 0;
    c.breakForce[i] = packedBreak(k.breakForce);
    c.broken[i] = k.broken ? 1 : 0;
  }
}

/** Full CPU state -> device (gpuworld.js:250-298). */
function upload(g) {
  const cpu = g.cpu;
  const b = g.body;
  for (let i = 0; i < g.nB; i++) {
    b.px[i] = cpu.px[i]; b.py[i] = cpu.py[i]; b.vx[i] = cpu.vx[i]; b.vy[i] = cpu.vy[i];
    b.ang[i] = cpu.ang[i]; b.av[i] = cpu.av[i];
    b.fx[i] = cpu.fx ? cpu.fx[i] : 0; b.fy[i] = cpu.fy ? cpu.fy[i] : 0; b.tq[i] = cpu.tq ? cpu.tq[i] : 0;
  }
  fillControls(g);
  const cs = cpu.constraints;
  const c = g.cons;
  for (let i = 0; i < g.nC; i++) {
    c.lambda[i] = cs[i].lambda || 0; c.lambdaX[i] = cs[i].lambdaX || 0; c.lambdaY[i] = cs[i].lambdaY || 0;
  }
  lib().upload(g.handle, b, c);
}

/** packControls (gpuworld.js:331-341), pure: the control fields of each
 * 80-byte constraint record — lbx o+6, lby o+7, compliance o+8, target o+10,
 * breakForce o+16 (inf -> 1e30), broken o+17 through the u32 view. */
function packControls(cpu, consData) {
  const u = new Uint32Array(consData.buffer, consData.byteOffset, consData.length);
  const cs = cpu.constraints;
  for (let i = 0; i < cs.length; i++) {
    const c = cs[i];
    const o = i * GPU_STRIDES.constraint;
    consData[o + 6] = c.lbx;
    consData[o + 7] = c.lby;
    consData[o + 8] = c.compliance;
    consData[o + 10] = c.target;
    consData[o + 16] = packedBreak(c.breakForce);
    u[o + 17] = c.broken ? 1 : 0;
  }
  return consData;
}

/** uploadControls (gpuworld.js:312-315): only the per-step control fields. */
function uploadControls(g) {
  packControls(g.cpu, g.consData);
  fillControls(g);
  lib().packControls(g.handle, g.cons);
}

/** One simulation step: S substeps + grip break (gpuworld.js:374-382). */
function step(g) {
  const cpu = g.cpu;
  let colors = 0;
  if (cpu.colorCnt) for (let c = 0; c < MAX_COLORS; c++) if (cpu.colorCnt[c] > 0) colors++;
  g.dispatches += 1 + cpu.substeps * (3 + cpu.iterations * (colors + 1));
  return lib().step(g.handle, 1);
}

function mirrorFields(cpu, i, c, lam, lamX, lamY, broken) {
  c.lambda = lam;
  c.broken = broken;
  if (c.type === C_GRIP) {
    c.lambdaX = lamX;
    c.lambdaY = lamY;
    c.force = Math.hypot(lamX, lamY) * (1 / cpu.h);
  }
}

/** One substep (gpuworld.js:344-372): lambda reset, integrate, the
 * iterations of colour passes and collisions, finalize; no grip break. */
function substep(g) {
  const cpu = g.cpu;
  let colors = 0;
  if (cpu.colorCnt) for (let c = 0; c < MAX_COLORS; c++) if (cpu.colorCnt[c] > 0) colors++;
  g.dispatches += 3 + cpu.iterations * (colors + 1);
  return lib().substep(g.handle, 1);
}

/** respawnCreature (src/sim/train.js:188-229) on the device, in place of a
 * host respawn and a full upload: creatures[i] goes to poses[3i..3i+2]
 * (x, y, heading); arms re-chained, velocities 0, motor targets 0, grips
 * released. The caller's world sees it at the next download. */
function respawn(g, creatures, poses) {
  const ids = creatures instanceof Int32Array ? creatures : Int32Array.from(creatures);
  const p = poses instanceof Float64Array ? poses : Float64Array.from(poses);
  lib().respawn(g.handle, ids, p);
}

/** Device -> the caller's world (gpuworld.js:385-396 + mirrorInto): bodies
 * px, py, vx, vy, ang, av; lambda and broken of every constraint; lambdaX,
 * lambdaY and force = |lambda|/h of grips. Returns the stats. */
async function download(g) {
  const cpu = g.cpu;
  const st = await lib().download(g.handle, g.body, g.cons);
  const b = g.body;
  for (let i = 0; i < g.nB; i++) {
    cpu.px[i] = b.px[i]; cpu.py[i] = b.py[i]; cpu.vx[i] = b.vx[i]; cpu.vy[i] = b.vy[i];
    cpu.ang[i] = b.ang[i]; cpu.av[i] = b.av[i];
  }
  const cs = cpu.constraints;
  const c = g.cons;
  for (let i = 0; i < g.nC; i++) {
    mirrorFields(cpu, i, cs[i], c.lambda[i], c.lambdaX[i], c.lambdaY[i], c.broken[i] !== 0);
  }
  const stats = { maxSpeed: st.maxSpeed, maxAngSpeed: st.maxAngSpeed, nan: st.nan };
  if (cpu.stats) Object.assign(cpu.stats, stats);
  return stats;
}

/** mirrorInto (gpuworld.js:417-456), pure: 48-byte body records (px o+0,
 * py o+1, vx o+4, vy o+5, ang o+6, av o+8) and 80-byte constraint records
 * (lambdaX o+13, lambdaY o+14, lambda o+15, broken o+17 as u32) into the
 * world; stats from the mirrored bodies. */
function mirrorInto(cpu, bodyData, consData) {
  const B = GPU_STRIDES.body;
  const C = GPU_STRIDES.constraint;
  let maxS2 = 0;
  let maxA = 0;
  let nan = 0;
  for (let i = 0; i < cpu.n; i++) {
    const o = i * B;
    const px = bodyData[o], py = bodyData[o + 1], vx = bodyData[o + 4], vy = bodyData[o + 5];
    const ang = bodyData[o + 6], av = bodyData[o + 8];
    cpu.px[i] = px; cpu.py[i] = py; cpu.vx[i] = vx; cpu.vy[i] = vy; cpu.ang[i] = ang; cpu.av[i] = av;
    const s2 = vx * vx + vy * vy;
    if (s2 > maxS2) maxS2 = s2;
    const aa = Math.abs(av);
    if (aa > maxA) maxA = aa;
    if (!(Number.isFinite(px) && Number.isFinite(py) && Number.isFinite(vx) && Number.isFinite(vy) &&
          Number.isFinite(ang))) nan++;
  }
  const u = new Uint32Array(consData.buffer, consData.byteOffset, consData.length);
  const cs = cpu.constraints;
  for (let i = 0; i < cs.length; i++) {
    const o = i * C;
    mirrorFields(cpu, i, cs[i], consData[o + 15], consData[o + 13], consData[o + 14], u[o + 17] !== 0);
  }
  const stats = { maxSpeed: Math.sqrt(maxS2), maxAngSpeed: maxA, nan };
  if (cpu.stats) Object.assign(cpu.stats, stats);
  return stats;
}

/** groupCollPoints (gpuworld.js:189-207), pure: points grouped by body with a
 * stable counting sort; range = [offsets of n bodies, counts of n bodies].
 * Throws on a point whose body is outside [0, n). */
function groupCollPoints(collPts, n) {
  const np = collPts.length;
  const cnt = new Uint32Array(n);
  for (let i = 0; i < np; i++) {
    const b = collPts[i].body;
    if (!(b >= 0 && b < n)) throw new Error(`groupCollPoints: point ${i} has body ${b} outside [0, ${n})`);
    cnt[b]++;
  }
  const range = new Uint32Array(2 * n);
  let acc = 0;
  for (let b = 0; b < n; b++) {
    range[b] = acc;
    range[n + b] = cnt[b];
    acc += cnt[b];
  }
  const fill = new Uint32Array(n);
  const order = new Uint32Array(np);
  for (let i = 0; i < np; i++) {
    const b = collPts[i].body;
    order[range[b] + fill[b]++] = i;
  }
  return { order, range };
}

/** Idempotent release of the device world (gpuworld.js:466-472). */
function destroy(g) {
  if (!g || !g.handle) return;
  try {
    lib().destroy(g.handle);
  } catch (e) {
    // already released
  }
  g.handle = null;
}

/** K steps under the device-side control law (no host round trip); events,
 * when given, is a Uint32Array of nsteps x creatures event words. */
function rollout(g, law, step0, nsteps, events) {
  return lib().rollout(g.handle, law || {}, step0 | 0, nsteps | 0, events);
}

/** Per-creature state readout in place of a full download: readProprio's
 * 6A+6 inputs (src/sim/core.js:129-175) then the mantle state (12 floats),
 * `stride` floats per creature into `out` (Float32Array of creatures x
 * stride). Returns the floats per creature the readout writes. */
function readState(g, out, stride) {
  return lib().readState(g.handle, { out }, stride | 0);
}

module.exports = {
  GPU_STRIDES, C_PIN, C_ANGLE, C_GRIP,
  isAvailable, create, upload, uploadControls, packControls, step, substep, download, mirrorInto,
  groupCollPoints, destroy, rollout, readState, respawn,
};
