'use strict';
// Device-side face of src/brain/net.js and src/brain/es.js, as
// src/sim/gputrain.js:55-86 and src/sim/train.js:154-158 drive them: the
// population's genomes stay on the GPU, one forwardInto call runs every
// creature, beginGeneration writes a whole generation of genomeAt genomes,
// update runs es.js's Adam step. Every call goes to libkxpbd (kx_policy_*,
// kx_es_*) through kxpbd.node; a negative return code throws as in gpuworld.js.
const path = require('path');

let native = null;
function lib() {
  if (!native) native = require(path.join(__dirname, 'kxpbd.node'));
  return native;
}

const HIDDEN = 24;               // src/brain/net.js:20
const OUT_DIM = 1 + 6 * 3 + 1;   // src/brain/net.js:23

function inputDim(proprioDim, senseDim) {
  return proprioDim + senseDim + OUT_DIM;
}

function weightCount(nIn, hidden = HIDDEN, nOut = OUT_DIM) {
  return lib().policyWeights(nIn, hidden, nOut);
}

/** n creatures' genomes (randomGenome's layout, one flat Float32Array of
 * n * weightCount(nIn) as gputrain.js:55 holds them) resident on the GPU. */
function createPopulation(n, nIn, hidden = HIDDEN, nOut = OUT_DIM) {
  return { handle: lib().policyCreate(n, nIn, hidden, nOut), n, nIn, hidden, nOut,
    nW: weightCount(nIn, hidden, nOut) };
}

function setGenomes(pop, genomes) {
  lib().policyGenomes(pop.handle, { genomes }, true, pop.n, pop.nW);
}

function getGenomes(pop, out) {
  const g = out || new Float32Array(pop.n * pop.nW);
  lib().policyGenomes(pop.handle, { genomes: g }, false, pop.n, pop.nW);
  return g;
}

/** forwardInto (net.js:71-98) for every creature: inputs [n * nIn] (the
 * previous outputs already in the last OUT_DIM inputs, gputrain.js:82),
 * outputs [n * nOut]; opts.scratch [n * 2 * hidden], opts.controls
 * [n * 25] (decode, net.js:114-124, of opts.arms arms as control-tape
 * records). Returns the kernel's device time in ms. */
function forwardInto(pop, inputs, outputs, opts) {
  const o = opts || {};
  return lib().policyForward(pop.handle, { x: inputs, out: outputs, scratch: o.scratch, controls: o.controls },
    o.arms | 0, pop.n, pop.nIn, pop.nOut, pop.hidden);
}

/** createES (es.js:22-58): theta zeros unless given; opts use es.js's names. */
function createES(nWeights, opts, theta) {
  const o = opts || {};
  let popSize = o.popSize === undefined ? 64 : o.popSize;
  if (popSize % 2 !== 0) popSize++;
  return { handle: lib().esCreate(nWeights, o, { theta }), nWeights, popSize };
}

/** beginGeneration's genomeAt loop (train.js:154-158) into pop's genomes,
 * from the trainer's stream state; returns the state after the draws. */
function beginGeneration(es, pop, rndState) {
  return lib().esGeneration(es.handle, pop.handle, rndState >>> 0);
}

/** update (es.js:92-149) -> { best, mean, sd, sigma } */
function update(es, fitness) {
  const f = fitness instanceof Float64Array ? fitness : Float64Array.from(fitness);
  return lib().esUpdate(es.handle, { fitness: f }, es.popSize);
}

/** theta, m, v, bestTheta (Float32Array) and sigma, t, bestEver */
function esState(es) {
  const n = es.nWeights;
  const out = { theta: new Float32Array(n), m: new Float32Array(n), v: new Float32Array(n),
    bestTheta: new Float32Array(n) };
  return Object.assign(out, lib().esGet(es.handle, out, n));
}

/** State of the trainer's stream before its first call (train.js:65). */
function trainerStreamState(seed) {
  return (seed * 2654435761) >>> 0;
}

function destroy(obj) {
  if (obj && obj.handle) lib().brainDestroy(obj.handle);
}

module.exports = {
  HIDDEN, OUT_DIM, inputDim, weightCount, createPopulation, setGenomes, getGenomes, forwardInto,
  createES, beginGeneration, update, esState, trainerStreamState, destroy,
};
