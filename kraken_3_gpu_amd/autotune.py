"""Autotuner for the persistent path's launch shape.

`tune` times short device rollouts for every (lanes per creature G,
wavefronts per workgroup) candidate that fits in LDS and keeps the fastest;
results are cached per morphology signature so a serving loop tunes each
bucket once.

`tune_buckets` picks a lane count per morphology bucket for one mixed launch
(kx_set_template_lanes): templates are grouped into size buckets and the
bucket lane counts are tuned together on the combined batch, since each one
changes both its own waves' work and the whole launch's LDS size. The steps
per launch (the rollout tile) and the wavefronts per workgroup are then
chosen on the same batch. Lane counts never change results, only speed.
"""
import json
import os
import time

import numpy as np

from . import _lib as K
from . import gpuworld as G

SHAPES = [(8, 1), (8, 2), (16, 1), (16, 2), (16, 4), (32, 1), (32, 2), (64, 1), (64, 2)]
LANES = (8, 16, 32, 64)
_cache = {}


def signature(g):
    return f"nCr={g.nCr}|nB={g.nB}|nC={g.nC}|tmpl={len(G.schedule_rounds(g))}"


def tune(g, law, steps=8, reps=2, shapes=SHAPES, cache_path=None):
    """Returns ((lanes, waves_per_block), {shape: ms}) and leaves g on the
    best shape. Device state advances by len(shapes)*reps*steps steps, so
    callers tune on a scratch handle or re-upload afterwards."""
    key = signature(g)
    if cache_path and os.path.exists(cache_path):
        try:
            _cache.update(json.load(open(cache_path)))
        except (OSError, ValueError):
            pass
    if key in _cache:
        best = tuple(_cache[key]["best"])
        G.set_path(g, K.KX_PATH_PERSISTENT, *best)
        return best, _cache[key]["times"]
    times = {}
    step = 0
    for shape in shapes:
        try:
            G.set_path(g, K.KX_PATH_PERSISTENT, *shape)
        except K.KxError:
            continue
        G.rollout(g, law, step, steps)            # first launch pays setup
        step += steps
        best_ms = None
        for _ in range(reps):
            G.rollout(g, law, step, steps)
            step += steps
            ms, _ = G.last_timing(g)
            best_ms = ms if best_ms is None else min(best_ms, ms)
        times[f"{shape[0]}x{shape[1]}"] = best_ms
    best = min(times, key=times.get)
    shape = tuple(int(x) for x in best.split("x"))
    G.set_path(g, K.KX_PATH_PERSISTENT, *shape)
    _cache[key] = {"best": list(shape), "times": times}
    if cache_path:
        with open(cache_path, "w") as f:
            json.dump(_cache, f)
    return shape, times


def _time_rollout(g, law, steps, reps):
    G.rollout(g, law, 0, steps)
    best = None
    for r in range(reps):
        G.rollout(g, law, steps * (r + 1), steps)
        ms, _ = G.last_timing(g)
        best = ms if best is None else min(best, ms)
    return best


def bucket_of_templates(tinfo, n_buckets):
    """Template -> bucket index: size classes by constraints per creature
    (population-weighted quantiles, so every bucket holds a similar share)."""
    nc = np.array([t["constraints"] for t in tinfo], dtype=np.float64)
    pop = np.array([t["creatures"] for t in tinfo], dtype=np.float64)
    order = np.argsort(nc, kind="stable")
    cum = np.cumsum(pop[order]) / max(pop.sum(), 1.0)
    b = np.minimum((cum * n_buckets - 1e-9).astype(int), n_buckets - 1)
    out = np.zeros(len(tinfo), dtype=int)
    out[order] = np.maximum(b, 0)
    return out


def tune_buckets(w, law, n_buckets=4, lanes=(8, 16, 32, 64), wpbs=(1, 2), tiles=(25, 50, 100),
                 steps=8, reps=2, passes=2):
    """Per-bucket lane counts by coordinate descent on the whole batch (one
    bucket's lane count changed at a time, kept when the launch gets faster),
    then (waves per workgroup, steps per launch). Lane counts move both the
    per-wave work and the launch's LDS size (the largest bucket region sets
    it), so they are timed together on the real mixed launch. Returns
    (handle, report); the handle holds w, uploaded, on the tuned shape."""
    t0 = time.perf_counter()
    g = G.create(w, path=K.KX_PATH_PERSISTENT, lanes_per_creature=16, waves_per_block=1)
    G.upload(g)
    tinfo = G.template_info(g)
    bucket = bucket_of_templates(tinfo, n_buckets)
    nb = int(bucket.max()) + 1
    trials = {}

    def per_template(bl):
        return [int(bl[bucket[t]]) for t in range(len(tinfo))]

    def measure(bl):
        key = "-".join(str(x) for x in bl)
        if key not in trials:
            try:
                G.set_template_lanes(g, per_template(bl))
            except K.KxError:
                trials[key] = None
                return None
            trials[key] = _time_rollout(g, law, steps, reps)
        return trials[key]

    cur = [16] * nb
    best_ms = measure(cur)
    for _ in range(passes):
        improved = False
        for b in range(nb):
            for G_ in lanes:
                if G_ == cur[b]:
                    continue
                cand = list(cur)
                cand[b] = G_
                ms = measure(cand)
                if ms is not None and ms < best_ms * 0.99:
                    cur, best_ms, improved = cand, ms, True
        if not improved:
            break
    G.set_template_lanes(g, per_template(cur))
    report = {"templates": len(tinfo), "buckets": [], "lane_trials_ms_8steps": trials}
    for b in range(nb):
        ts = [t for t in range(len(tinfo)) if bucket[t] == b]
        report["buckets"].append({"templates": len(ts), "creatures": int(sum(tinfo[t]["creatures"] for t in ts)),
                                  "constraints": [int(min(tinfo[t]["constraints"] for t in ts)),
                                                  int(max(tinfo[t]["constraints"] for t in ts))],
                                  "lanes": int(cur[b])})
    best_shape, best_rate, shape_trials = None, 0.0, {}
    for wpb in wpbs:
        try:
            G.set_path(g, K.KX_PATH_PERSISTENT, 0, wpb)
        except K.KxError:
            continue
        for tile in tiles:
            ms = _time_rollout(g, law, tile, reps)
            rate = g.nCr * tile / (ms * 1e-3)
            shape_trials[f"wpb{wpb}_tile{tile}"] = rate
            if rate > best_rate:
                best_shape, best_rate = (wpb, tile), rate
    G.set_path(g, K.KX_PATH_PERSISTENT, 0, best_shape[0])
    G.upload(g)
    report.update(waves_per_block=best_shape[0], steps_per_launch=best_shape[1], shape_trials=shape_trials,
                  seconds=time.perf_counter() - t0, launch_shape=G.persist_info(g))
    return g, report
