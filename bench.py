"""Throughput benchmark: batched XPBD creature steps on one MI355X.

Workload (uniform batch): 4096 identical creatures (mantle + 6 arms x 7
segments: 43 bodies, 90 constraints, 92 collision points each) on
makeMap(1200, 800, seed 1), +-20 px spawn jitter, S=4 substeps x I=8
iterations, h = 1/240 s, device-side seeded control law with the adhesion
reflex. One bench step = one simulation step of the whole batch.

value = creature-steps/s of the persistent batched path (state resident in
HBM when the timed region starts); the plain per-colour-launch path runs the
same workload beside it. Prints one JSON line.

  python bench.py [--gpus 1] [--steps 1000] [--warmup 20]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3       # MI355X published FP32 vector peak (packed FMA)
FP32_UNPACKED_TFLOPS = 78.6    # the same without packed-FP32 instructions, which this build does not emit
HBM_PEAK_GBS = 8000.0


def flops_per_creature_step(S, I, pins, motors, bodies, points, grips_attached, rects_near, grips_total):
    """Algorithmic FLOPs of one creature-step (per-op counts of the WGSL
    bodies: pin 85, motor 42, attached grip 66, point transform 11, rect
    test 12, integrate+finalize 45 per body, grip break 5 per grip)."""
    it = 85 * pins + 42 * motors + 66 * grips_attached + points * (11 + 12 * rects_near)
    return S * (I * it + 45 * bodies) + 5 * grips_total


def rects_within(w, radius=16.0):
    """Mean number of rects within `radius` px of each collision point."""
    b = w.p_body
    ca, sa = np.cos(w.ang[b]), np.sin(w.ang[b])
    x = w.px[b] + w.p_lx * ca - w.p_ly * sa
    y = w.py[b] + w.p_lx * sa + w.p_ly * ca
    r = w.rects.astype(np.float64)
    cnt = np.zeros(len(x))
    for rx, ry, rw, rh in r:
        dx = x - np.clip(x, rx, rx + rw)
        dy = y - np.clip(y, ry, ry + rh)
        cnt += (dx * dx + dy * dy) <= radius * radius
    return float(cnt.mean())


def algorithmic_launch_bytes(w, nsteps):
    """HBM bytes one persistent launch must move: the state it reads at entry
    and writes at exit (the LDS holds everything in between) plus one event
    word per creature and step.
    in:  bodies px, py, ang, vx, vy, av (6 f32); constraints compliance,
         target, lbx, lby, breakForce (5 f32) + broken (u8); 8 f64 reflex
         cooldowns per creature
    out: bodies 9 f32 (the 6 above + predicted ppx, ppy, angP); constraints
         broken, breakForce, target, lambda, lambda_x, lambda_y (1 B + 5 f32),
         compliance of motors, world anchors of world-anchored constraints;
         cooldowns; events"""
    from kraken_3_gpu_amd import _lib as K
    nb, nc, ncr = w.n, w.nc, w.n_creatures
    n_ang = int((w.type == K.KX_ANGLE).sum())
    n_world = int((w.b < 0).sum())
    bin_ = nb * 6 * 4 + nc * (5 * 4 + 1) + ncr * 8 * 8
    bout = nb * 9 * 4 + nc * (1 + 5 * 4) + n_ang * 4 + n_world * 8 + ncr * 8 * 8 + ncr * 4 * nsteps
    return bin_ + bout


def measured_traffic(n_cr, chunk, workload="uniform4096", kernel="kx_k_persist"):
    """HBM bytes per launch from the committed rocprofv3 counter summaries
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), when one covers this
    kernel, workload and launch length."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"{kernel}_{workload}*_traffic.json"))):
        try:
            with open(path) as f:
                t = json.load(f)
        except (OSError, ValueError):
            continue
        if t.get("creatures") == n_cr and t.get("steps_per_launch") == chunk:
            return dict(t, source=os.path.relpath(path, ROOT))
    return None


def state_readout_cost(g, w, reps=20):
    """What a host policy loop pays per step to see the population: the
    per-creature readout (kx_read_state: readProprio inputs + mantle state)
    against the full state download the reference's loop mirrors (bodies and
    constraints), both timed on the host around the call (device work, the
    copy over PCIe and the synchronisation included)."""
    def timed(fn):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) * 1e3 / reps
    from kraken_3_gpu_amd import gpuworld as G
    st = G.read_state(g)
    r_ms = timed(lambda: G.read_state(g))
    d_ms = timed(lambda: G.download(g))
    nb, nc = len(w.px), len(w.a)
    return {"readout_ms": r_ms, "readout_bytes": int(st.nbytes), "floats_per_creature": int(st.shape[1]),
            "download_ms": d_ms, "download_bytes": int(nb * 24 + nc * 13)}


def policy_cost(n, reps=50):
    """The policy network and ES over a population of n creatures with the
    reference's network (105 -> 24 -> 24 -> 20, one genome per creature): the
    forward kernel's device time over resident genomes (HIP events, mean of
    `reps` launches) against its algorithmic HBM bytes (genomes + inputs +
    outputs), the host-inclusive forward (inputs in, outputs and decoded
    control records out), and one ES generation (genomeAt for every member)
    and update."""
    from kraken_3_gpu_amd import brain as B
    n_in = B.input_dim()
    rng = np.random.default_rng(3)
    nw = B.weight_count(n_in)
    res = {}
    with B.Population(n, n_in) as P:
        P.set_genomes((rng.standard_normal((n, nw)) / np.sqrt(n_in)).astype(np.float32))
        x = rng.uniform(-1, 1, (n, n_in)).astype(np.float32)
        P.forward_into(x, arms=6)
        k_ms = P.forward_repeat(reps)
        t0 = time.perf_counter()
        for _ in range(reps):
            P.forward_into(x, arms=6)
        host_ms = (time.perf_counter() - t0) * 1e3 / reps
        alg = n * (nw + n_in + B.OUT_DIM) * 4
        res["forward"] = {"kernel": "kx_k_policy_fwd", "creatures": n, "weights_per_creature": nw,
                          "kernel_ms": k_ms, "algorithmic_bytes": alg, "achieved_gbs": alg / (k_ms * 1e6),
                          "frac_peak": alg / (k_ms * 1e6) / HBM_PEAK_GBS, "host_inclusive_ms": host_ms,
                          "working_set": "genomes re-read from the 256 MB Infinity Cache (%.1f MB resident)"
                                         % (alg / 1e6)}
    # the same kernel over a population whose genomes (358 MB) exceed the
    # Infinity Cache, so every launch streams them from HBM
    nbig = 24576
    with B.Population(nbig, n_in) as P:
        P.set_genomes((rng.standard_normal((nbig, nw)) / np.sqrt(n_in)).astype(np.float32))
        x = rng.uniform(-1, 1, (nbig, n_in)).astype(np.float32)
        P.forward_into(x, arms=6)
        k_ms = P.forward_repeat(max(5, reps // 5))
        alg = nbig * (nw + n_in + B.OUT_DIM) * 4
        res["forward_hbm"] = {"kernel": "kx_k_policy_fwd", "creatures": nbig, "kernel_ms": k_ms,
                              "algorithmic_bytes": alg, "achieved_gbs": alg / (k_ms * 1e6),
                              "frac_hbm": alg / (k_ms * 1e6) / HBM_PEAK_GBS,
                              "working_set": "%.0f MB of genomes > the 256 MB Infinity Cache" % (alg / 1e6)}
    pe = min(n, 8192)                      # the ES's population limit
    with B.ES(nw, popSize=pe) as es, B.Population(pe, n_in) as PE:
        state = B.trainer_stream_state(1)
        es.begin_generation(PE, state)
        t0 = time.perf_counter()
        for _ in range(5):
            state = es.begin_generation(PE, state)
        gen_ms = (time.perf_counter() - t0) * 1e3 / 5
        fit = rng.normal(0, 10, es.pop_size)
        t0 = time.perf_counter()
        for _ in range(5):
            es.update(fit)
        upd_ms = (time.perf_counter() - t0) * 1e3 / 5
    res["es"] = {"pop": pe, "generation_ms": gen_ms, "update_ms": upd_ms,
                 "generation_bytes": pe * nw * 4 + (pe // 2) * nw * 4}
    return res


def policy_loop_cost(n, steps=40):
    """gputrain.js's per-step loop (readProprio -> forwardInto -> decode ->
    applyControl -> step) at n creatures with the reference's network and
    random senses, two ways: kx_policy_drive (everything but the senses on
    the device) and the host-driven loop over the same library (read_state
    down, inputs assembled on the host, forward_into, upload_tape, a one-step
    tape rollout). Host-timed per step, the stream drained at the end; the
    physics-only one-step rollout under the device law is timed beside them."""
    from kraken_3_gpu_amd import _lib as K
    from kraken_3_gpu_amd import brain as B
    from kraken_3_gpu_amd import gpuworld as G
    from kraken_3_gpu_amd import world as W
    n_in = B.input_dim()
    rng = np.random.default_rng(4)
    gen = (rng.standard_normal((n, B.weight_count(n_in))) / np.sqrt(n_in)).astype(np.float32)
    senses = rng.uniform(-1, 1, (n, B.SENSE_DIM)).astype(np.float32)
    law = W.default_law(5)
    tlaw = W.default_law(5)
    tlaw.flags |= K.KX_LAW_TAPE
    res = {"creatures": n, "steps": steps}
    w = W.world_uniform(n, seed=2)
    g = G.create(w)
    G.upload(g)
    with B.Population(n, n_in) as P:
        P.set_genomes(gen)
        for s in range(3):
            P.drive(g, law, s, senses)
        G.sync(g)
        t0 = time.perf_counter()
        for s in range(steps):
            P.drive(g, law, s, senses)
        G.sync(g)
        res["device_loop_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
        prev = np.zeros((n, B.OUT_DIM), np.float32)
        t0 = time.perf_counter()
        for s in range(steps):
            st = G.read_state(g)
            x = np.concatenate([st[:, :42], senses, prev], axis=1)
            prev, ctl = P.forward_into(x, arms=6)
            G.upload_tape(g, ctl[None])
            G.rollout(g, tlaw, s, 1)
        G.sync(g)
        res["host_loop_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
    t0 = time.perf_counter()
    for s in range(steps):
        G.rollout(g, law, s, 1)
    G.sync(g)
    res["physics_only_ms_per_step"] = (time.perf_counter() - t0) * 1e3 / steps
    G.destroy(g)
    res["device_loop_creature_steps_per_s"] = n / (res["device_loop_ms_per_step"] * 1e-3)
    return res


def device_parity(g, w, law, step, every=64):
    """The device's error against the f64 oracle on the benchmarked world: a
    creature subset (every `every`-th) pulled from the device after the timed
    run, one more step of the same law on both sides from the identical
    binary32 state. Returns max |dpos| (px) over creatures whose grip events
    agree, and the creatures whose events differ (flips). The oracle is the
    checker here, as in the parity tests."""
    import ctypes as C
    from oracle import pyoracle as po
    from kraken_3_gpu_amd import gpuworld as G
    G.download(g)
    ctl = G.download_controls(g)
    G.download_cooldown(g)
    f32 = lambda a: np.asarray(a, np.float32).astype(np.float64)
    sub = list(range(0, w.n_creatures, every))
    o = po.World(w.S, w.I, w.h)
    o.set_map(w.width, w.height, w.rects.astype(np.float64))
    lo = np.concatenate([[0], np.cumsum([1 + sp[3] * sp[4] for sp in w.spawns])])
    bidx = []
    for c in sub:
        x, y, a, arms, segs, ring, curl = w.spawns[c]
        o.spawn(x, y, a, arms, segs, ring, curl)
        bidx.extend(range(lo[c], lo[c + 1]))
    o.color()
    bidx = np.asarray(bidx)
    crb = np.searchsorted(lo, np.arange(w.n), side="right") - 1
    cidx = np.nonzero(np.isin(crb[w.a], sub))[0]
    o.set_bodies(*(f32(getattr(w, k)[bidx]) for k in ("px", "py", "ang", "vx", "vy", "av")))
    o.set_cons_dynamic(broken=w.broken[cidx].astype(np.int32), lbx=f32(ctl["lbx"][cidx]), lby=f32(ctl["lby"][cidx]),
                       compliance=f32(ctl["compliance"][cidx]), target=f32(ctl["target"][cidx]),
                       breakForce=f32(ctl["breakForce"][cidx]))
    o.set_cd(np.ascontiguousarray(w.cd[sub]))
    o.set_step_counter(step)
    o.clear_events()
    for local, key in enumerate(sub):
        po.lib().or_apply_law_one(o.p, C.c_uint32(law.seed), C.c_uint32(step), local, key)
    o.step()
    ev = G.rollout(g, law, step, 1, events=True)
    G.download(g)
    ev_gpu = set()
    for c in sub:
        v = int(ev[0, c])
        for bit in range(24):
            if v >> bit & 1:
                ev_gpu.add((sub.index(c), bit // 3, bit % 3))
    ev_or = {(c, a, k) for (_, c, a, k) in o.events()}
    flips = sorted({e[0] for e in ev_gpu ^ ev_or})
    b = o.bodies()
    err = np.maximum(np.abs(w.px[bidx] - b["px"]), np.abs(w.py[bidx] - b["py"]))
    local = np.searchsorted(lo[sub], bidx, side="right") - 1
    ok = ~np.isin(local, flips)
    return {"max_dpos_px": float(err[ok].max()), "max_dpos_px_all": float(err.max()), "event_flips": len(flips),
            "events": len(ev_or), "creatures_checked": len(sub), "step": step,
            "check": "one step of the device vs the f64 oracle from identical binary32 state"}


def cpu_baseline(seconds=12.0, device=None):
    """The oracle (C restatement of xpbd.js, f64) stepping 64-creature worlds
    under the same control law on the host cores, one world per thread; with
    `device` = (handle, world, law, step), also the device's one-step error
    against it on a subset of the benchmarked world (device_parity)."""
    from oracle import pyoracle as po
    from kraken_3_gpu_amd import world as W
    threads = max(1, min(16, os.cpu_count() or 1))
    w = W.world_uniform(64, seed=1)
    o = po.World(w.S, w.I, w.h)
    o.set_map(w.width, w.height, w.rects.astype(np.float64))
    for (x, y, a, arms, segs, ring, curl) in w.spawns:
        o.spawn(x, y, a, arms, segs, ring, curl)
    o.color()
    rate1 = o.bench_threads(1, 2, 5)           # calibration: creature-steps/s on one thread
    steps = max(2, int(seconds * rate1 / 64 / 1.0))
    rate = o.bench_threads(threads, steps, 5)
    out = {"value": rate, "unit": "creature-steps/s", "cores": threads, "kind": "port",
           "sample": f"{threads} threads x 64 creatures x {steps} steps (uniform batch recipe, f64 oracle)"}
    if device is not None:
        out["device_parity"] = device_parity(*device)
    return out


def copy_probe(mib):
    """Three plain device copies of a known size: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters to bytes moved."""
    import torch
    n = mib << 20
    x = torch.ones(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n, "copy_probe_bytes_written": n, "copies": 3}), flush=True)


def law_for(workload, seed):
    """The control law of a workload: the reference's reflex constants, or for
    the stress run every tip adhesive with a per-tip break limit drawn once
    from U[50, 200] N."""
    from kraken_3_gpu_amd import _lib as K
    from kraken_3_gpu_amd import world as W
    if workload == "stress":
        law = W.default_law(seed, grip_min=50.0, grip_max=200.0)
        law.attach_min, law.release_below = -1.0, -1.0
        law.flags = K.KX_LAW_FIXED_GRIP
        return law
    return W.default_law(seed)


def flops_per_step(w, g_att_total, r_near):
    """Algorithmic FLOPs of one step of the whole world (the per-creature
    formula summed over its creatures, any morphology)."""
    from kraken_3_gpu_amd import _lib as K
    pins = int((w.type == K.KX_PIN).sum())
    motors = int((w.type == K.KX_ANGLE).sum())
    grips = int((w.type == K.KX_GRIP).sum())
    it = 85 * pins + 42 * motors + 66 * g_att_total + len(w.p_body) * (11 + 12 * r_near)
    return w.S * (w.I * it + 45 * w.n) + 5 * grips


def persistent_run(g, law, steps, warmup, chunk, step0=0):
    """Warm up, then time `steps` simulation steps in launches of `chunk`
    steps (state resident in HBM). Returns timing and the next step key."""
    from kraken_3_gpu_amd import gpuworld as G
    step = step0
    while step < step0 + warmup:
        n = min(chunk, step0 + warmup - step)
        G.rollout(g, law, step, n)
        step += n
    G.sync(g)
    chunk_ms = []
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        G.rollout(g, law, step, n)
        ms, _ = G.last_timing(g)
        chunk_ms.append((ms, n))
        step += n
        done += n
    G.sync(g)
    wall = time.perf_counter() - t0
    full = [m for m, n in chunk_ms if n == chunk] or [chunk_ms[0][0]]
    lat = sorted(full)
    return dict(wall=wall, chunk_ms=chunk_ms, launch_ms=float(np.mean(full)), step=step,
                p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                dev_ms=sum(m for m, _ in chunk_ms))


def rollout_latency(g, law, step, chunk, launches):
    """Rollout latency: `launches` launches of `chunk` steps each, one at a
    time, timed on the host from submission to completion and on the device
    with HIP events. Returns percentiles of both and the next step key."""
    from kraken_3_gpu_amd import gpuworld as G
    host, dev = [], []
    G.sync(g)
    for _ in range(launches):
        t0 = time.perf_counter()
        G.rollout(g, law, step, chunk)
        G.sync(g)
        host.append((time.perf_counter() - t0) * 1e3)
        dev.append(G.last_timing(g)[0])
        step += chunk
    h, d = np.asarray(host), np.asarray(dev)
    return dict(samples=launches, steps_per_launch=chunk, host_p50_ms=float(np.percentile(h, 50)),
                host_p99_ms=float(np.percentile(h, 99)), device_p50_ms=float(np.percentile(d, 50)),
                device_p99_ms=float(np.percentile(d, 99)), device_mean_ms=float(d.mean()),
                device_min_ms=float(d.min()), device_max_ms=float(d.max()), step=step)


def serving_fixed_rate(n_req, rate, seed=7):
    """ES evaluation requests (16-512 antithetic pairs, 300-1500 steps, 4-8
    arms x 5-9 segments, ring 6-10) arriving as a Poisson stream at a fixed
    `rate` per second into the continuous-batching server: request latency
    (arrival -> result) percentiles and the aggregate creature-steps/s."""
    from kraken_3_gpu_amd import serve as S
    t, reqs = S.es_stream(n_req, rate=rate, seed=seed)
    with S.Server(S.default_opts(chunk_steps=50)) as srv:
        _, summ = S.run_stream(srv, t, reqs)
    mean_cs = float(np.mean([2 * r["pairs"] * r["steps"] for r in reqs]))
    return dict(requests=n_req, arrival_rate_per_s=rate, offered_creature_steps_per_s=rate * mean_cs,
                p50_request_ms=summ["p50_ms"], p99_request_ms=summ["p99_ms"], max_request_ms=summ["max_ms"],
                creature_steps_per_s=summ["creature_steps_per_s"], p99_admit_wait_ms=summ["p99_admit_wait_ms"],
                device_busy_frac=summ["device_busy_frac"], failed=summ["failed"], peak_active=summ["peak_active"])


def warm_accounting(g, w):
    """Attached grips and rects near the collision points, from the device state."""
    from kraken_3_gpu_amd import _lib as K
    from kraken_3_gpu_amd import gpuworld as G
    G.download(g)
    grips = w.type == K.KX_GRIP
    return float((w.broken[grips] == 0).sum()), rects_within(w)


def uneven_leg(steps=300):
    """A bounded run of the uneven batch as the default line reports it: 8192
    creatures of 125 shapes, per-bucket lanes from a short autotune, 50-step
    launches replayed from a captured hipGraph."""
    from kraken_3_gpu_amd import autotune as A
    from kraken_3_gpu_amd import gpuworld as G
    from kraken_3_gpu_amd import world as W
    tile = 50
    w = W.world_uneven(8192, seed=1)
    law = law_for("uneven", 2024)
    g, rep = A.tune_buckets(w, law, steps=4, reps=1, passes=1, tiles=(tile,))
    G.upload(g)
    G.rollout_graph(g, law, 0, tile, 1)
    G.sync(g)
    t0 = time.perf_counter()
    G.rollout_graph(g, law, tile, tile, steps // tile)
    G.sync(g)
    wall = time.perf_counter() - t0
    dev_ms, launches = G.last_timing(g)
    st = G.download(g)
    out = {"workload": "uneven batch: 8192 creatures, 4-8 arms x 5-9 segments, ring 6-10, makeMap 1200x800",
           "creatures": w.n_creatures, "steps": steps, "steps_per_launch": tile,
           "value": w.n_creatures * steps / wall, "unit": "creature-steps/s", "launch_ms": dev_ms / launches,
           "lanes_per_bucket": [b["lanes"] for b in rep["buckets"]], "waves_per_block": rep["waves_per_block"],
           "final_stats": st}
    G.destroy(g)
    return out


def stress_leg(steps=200, warmup=20):
    """A bounded run of the stress batch as the default line reports it:
    65536 creatures on makeMap(4096, 4096, 32 x 32 corridors), every tip
    adhesive with a per-tip break limit from U[50, 200] N, 50-step launches."""
    from kraken_3_gpu_amd import _lib as K
    from kraken_3_gpu_amd import gpuworld as G
    from kraken_3_gpu_amd import world as W
    w = W.world_stress(65536, seed=1)
    law = law_for("stress", 2024)
    g = G.create(w, path=K.KX_PATH_PERSISTENT, waves_per_block=2)
    G.upload(g)
    r = persistent_run(g, law, steps, warmup, 50)
    st = G.download(g)
    out = {"workload": "stress: 65536 creatures (6 arms x 7 segments), makeMap(4096, 4096, 32 x 32 corridors)",
           "creatures": w.n_creatures, "steps": steps, "steps_per_launch": 50,
           "value": w.n_creatures * steps / r["wall"], "unit": "creature-steps/s", "launch_ms": r["launch_ms"],
           "launch_shape": G.persist_info(g), "final_stats": st}
    G.destroy(g)
    return out


def bench_stress(args):
    """65 536 creatures on makeMap(4096, 4096, 32 x 32 corridors), every tip
    adhesive with a per-tip break limit from U[50, 200] N."""
    from kraken_3_gpu_amd import _lib as K
    from kraken_3_gpu_amd import gpuworld as G
    from kraken_3_gpu_amd import world as W
    steps = args.steps or 2000
    law = law_for("stress", 2024)
    w = W.world_stress(args.creatures or 65536, seed=1)
    n_cr = w.n_creatures
    # two waves per workgroup: the template and bitmaps are shared by 8
    # creatures, so 4 workgroups (8 waves) fit a CU's LDS instead of 7
    g = G.create(w, path=K.KX_PATH_PERSISTENT, lanes_per_creature=args.lanes, waves_per_block=args.wpb or 2)
    G.upload(g)
    G.rollout(g, law, 0, args.warmup)
    G.sync(g)
    g_att, r_near = warm_accounting(g, w)
    r = persistent_run(g, law, steps, 0, args.chunk, step0=args.warmup)
    st = G.download(g)
    F = flops_per_step(w, g_att, min(r_near, 64.0))
    ach = F * args.chunk / (r["launch_ms"] * 1e-3) / 1e12
    lb = algorithmic_launch_bytes(w, args.chunk)
    shape = G.persist_info(g)
    kernel = ("kx_k_persist_reg" if shape.get("register_rounds") else
              "kx_k_persist_fix" if shape.get("fixed_rounds") else "kx_k_persist")
    traffic = measured_traffic(n_cr, args.chunk, "stress65536", kernel)
    out = {
        "metric": "creature-steps/s, stress run (65536 creatures, 4096x4096 map, 32x32 corridors)",
        "value": n_cr * steps / r["wall"], "unit": "creature-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": r["wall"] * 1e3 / steps, "higher_is_better": True,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"stress: {n_cr} creatures (6 arms x 7 segments) in the 32x32 corridor cells of "
                               f"makeMap(4096, 4096), all tips adhesive, per-tip break limit U[50,200] N",
                   "creatures": n_cr, "rects": int(len(w.rects)), "steps_per_launch": args.chunk},
        "roofline": {"bound": "valu", "achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / FP32_PEAK_TFLOPS, "launch_ms": r["launch_ms"],
                     "hbm_algorithmic_gbs": lb / (r["launch_ms"] * 1e-3) / 1e9, "hbm_peak_gbs": HBM_PEAK_GBS,
                     "traffic": traffic["traffic_bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "hbm_measured_gbs": (traffic["traffic_bytes"] / (r["launch_ms"] * 1e-3) / 1e9) if traffic else None,
                     "grips_attached": g_att, "rects_within_16px_mean": r_near},
        "launch_shape": shape, "kernel": kernel,
        "p50_rollout_ms": r["p50"], "p99_rollout_ms": r["p99"], "final_stats": st,
    }
    print(json.dumps(out), flush=True)


def bench_uneven(args):
    """8192 creatures of seeded morphology (4-8 arms x 5-9 segments, mantle
    ring 6-10) in one persistent launch: lanes per creature tuned per
    morphology bucket, waves per workgroup and steps per launch tuned on the
    whole batch, then 1500 steps replayed from a captured hipGraph."""
    from kraken_3_gpu_amd import _lib as K
    from kraken_3_gpu_amd import autotune as A
    from kraken_3_gpu_amd import gpuworld as G
    from kraken_3_gpu_amd import world as W
    steps = args.steps or 1500
    law = law_for("uneven", 2024)
    w = W.world_uneven(args.creatures or 8192, seed=1)
    n_cr = w.n_creatures
    g, rep = A.tune_buckets(w, law, tiles=(25, 50, 100) if not args.chunk_set else (args.chunk,))
    tile = rep["steps_per_launch"]
    steps = -(-steps // tile) * tile
    G.upload(g)
    G.rollout_graph(g, law, 0, tile, max(1, args.warmup // tile))     # capture + warm
    G.sync(g)
    g_att, r_near = warm_accounting(g, w)
    warm = max(1, args.warmup // tile) * tile
    G.sync(g)
    t0 = time.perf_counter()
    G.rollout_graph(g, law, warm, tile, steps // tile)
    G.sync(g)
    wall = time.perf_counter() - t0
    dev_ms, launches = G.last_timing(g)
    st = G.download(g)
    launch_ms = dev_ms / launches
    F = flops_per_step(w, g_att, min(r_near, 64.0))
    ach = F * tile / (launch_ms * 1e-3) / 1e12
    lb = algorithmic_launch_bytes(w, tile)
    value = n_cr * steps / wall
    plain = None
    if not args.no_plain:
        gp = G.create(w, path=K.KX_PATH_PLAIN)
        G.set_plain_sdf(gp, True)
        G.upload(gp)
        G.rollout(gp, law, 0, 2)
        G.sync(gp)
        t1 = time.perf_counter()
        G.rollout(gp, law, 2, args.plain_steps)
        G.sync(gp)
        pwall = time.perf_counter() - t1
        plain = {"value": n_cr * args.plain_steps / pwall, "steps": args.plain_steps, "sdf": "grid"}
        G.destroy(gp)
    sizes = np.array([a * s for a, s, _ in w.morph])
    out = {
        "metric": "creature-steps/s, uneven batch in one hipGraph-captured persistent launch",
        "value": value, "unit": "creature-steps/s", "n_gpus": 1, "steps": steps, "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"uneven batch: {n_cr} creatures, seeded morphology 4-8 arms x 5-9 segments, "
                               f"mantle ring 6-10, makeMap 1200x800 seed 1, {steps} steps from a hipGraph",
                   "creatures": n_cr, "bodies": int(w.n), "constraints": int(len(w.type)),
                   "arm_segments_range": [int(sizes.min()), int(sizes.max())], "steps_per_launch": tile},
        "roofline": {"bound": "valu", "achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / FP32_PEAK_TFLOPS, "launch_ms": launch_ms,
                     "hbm_algorithmic_gbs": lb / (launch_ms * 1e-3) / 1e9, "hbm_peak_gbs": HBM_PEAK_GBS},
        "autotune": rep, "launch_shape": G.persist_info(g), "final_stats": st,
    }
    if plain:
        out["plain"] = plain
        out["speedup_vs_plain"] = value / plain["value"]
    print(json.dumps(out), flush=True)


def bench_serve(args):
    """ES rollout requests through the continuous-batching server: a closed
    loop with 48 requests always outstanding first measures the serving
    capacity (creature-steps/s executed with the queue never empty), then a
    Poisson stream at 70 % of that capacity gives p50 / p99 request latency
    (arrival -> result) and the aggregate rate."""
    from kraken_3_gpu_amd import serve as S
    chunk = args.chunk if args.chunk_set else 50
    opts = S.default_opts(chunk_steps=chunk)
    n_req = args.requests
    with S.Server(opts) as srv:
        _, reqs = S.es_stream(64, rate=0.0, seed=101)
        probe = S.saturate(srv, reqs)
    cap = probe["creature_steps_per_s"]
    t, reqs = S.es_stream(n_req, rate=1.0, seed=7)
    mean_cs = float(np.mean([2 * r["pairs"] * r["steps"] for r in reqs]))
    rate = args.load * cap / mean_cs
    t = t / rate                                   # unit-rate gaps scaled to the target rate
    with S.Server(opts) as srv:
        res, summ = S.run_stream(srv, t, reqs)
    out = {
        "metric": "creature-steps/s of ES rollout requests served by continuous batching; p50/p99 request ms",
        "value": summ["creature_steps_per_s"], "unit": "creature-steps/s", "n_gpus": 1, "steps": None,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"serving: {n_req} ES requests, Poisson arrivals at {rate:.2f}/s "
                               f"({args.load:.0%} of the measured capacity), 16-512 antithetic pairs, "
                               f"300-1500 steps, morphology 4-8 arms x 5-9 segments, ring 6-10, "
                               f"controls from a 64-32-48 tanh policy on the host",
                   "requests": n_req, "arrival_rate_per_s": rate, "offered_load": args.load,
                   "mean_request_creature_steps": mean_cs, "steps_per_launch": chunk},
        "p50_request_ms": summ["p50_ms"], "p99_request_ms": summ["p99_ms"], "max_request_ms": summ["max_ms"],
        "p50_admit_wait_ms": summ["p50_admit_wait_ms"], "p99_admit_wait_ms": summ["p99_admit_wait_ms"],
        "capacity_creature_steps_per_s": cap, "capacity_probe": probe, "stream": summ,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["uniform", "uneven", "serve", "stress", "policy"], default="uniform",
                    help="uniform batch (the headline line), the uneven batch, the serving loop or the stress run")
    ap.add_argument("--requests", type=int, default=120, help="serve: requests in the Poisson stream")
    ap.add_argument("--load", type=float, default=0.7, help="serve: offered load as a fraction of capacity")
    ap.add_argument("--steps", type=int, default=0, help="timed steps (default 1000; stress 2000)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--creatures", type=int, default=0, help="creatures (default 4096; stress 65536)")
    ap.add_argument("--chunk", type=int, default=0,
                    help="simulation steps per persistent launch (default 50; uneven: tuned)")
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--latency-launches", type=int, default=100,
                    help="launches timed one by one for the rollout latency and the roofline's launch duration")
    ap.add_argument("--serve-requests", type=int, default=100,
                    help="requests of the fixed-rate serving leg (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="arrivals per second of the serving leg")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per creature (0 = default)")
    ap.add_argument("--wpb", type=int, default=0, help="wavefronts per workgroup (0 = default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the bounded uneven and stress legs")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    args.chunk_set = args.chunk > 0
    args.chunk = args.chunk or 50
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (creatures shard across "
                                   "GPUs with no per-step exchange; one process per GPU, each with its own share)"}))
        sys.exit(2)

    if args.workload == "stress":
        bench_stress(args)
        return
    if args.workload == "uneven":
        bench_uneven(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return
    if args.workload == "policy":
        n = args.creatures or 4096
        print(json.dumps({"workload": "policy network + ES", **policy_cost(n), "loop": policy_loop_cost(n)}),
              flush=True)
        return
    args.steps = args.steps or 1000
    args.creatures = args.creatures or 4096

    from kraken_3_gpu_amd import _lib as K
    from kraken_3_gpu_amd import gpuworld as G
    from kraken_3_gpu_amd import world as W

    seed = 2024
    law = W.default_law(seed)
    w = W.world_uniform(args.creatures, seed=1)
    n_cr = w.n_creatures
    # every launch runs `chunk` steps (the last one of the timed region may run fewer)
    chunk = min(args.chunk, args.steps)
    g = G.create(w, path=K.KX_PATH_PERSISTENT, lanes_per_creature=args.lanes, waves_per_block=args.wpb)
    G.upload(g)
    step = 0
    while step < args.warmup:
        n = min(chunk, args.warmup - step)
        G.rollout(g, law, step, n)
        step += n
    G.sync(g)
    # algorithmic accounting from the warmed-up state
    G.download(g)
    grips = w.type == K.KX_GRIP
    g_att = float((w.broken[grips] == 0).sum()) / n_cr
    r_near = rects_within(w)

    chunk_ms = []
    G.sync(g)
    t0 = time.perf_counter()
    done = 0
    while done < args.steps:
        n = min(chunk, args.steps - done)
        G.rollout(g, law, step, n)
        ms, _ = G.last_timing(g)
        chunk_ms.append((ms, n))
        step += n
        done += n
    G.sync(g)
    wall = time.perf_counter() - t0
    value = n_cr * args.steps / wall
    dev_ms = sum(m for m, _ in chunk_ms)
    st = G.download(g)
    shape = G.persist_info(g)
    # launch duration and rollout latency over many launches of the same length
    lat = rollout_latency(g, law, step, chunk, args.latency_launches)
    step = lat.pop("step")
    readout = state_readout_cost(g, w)
    policy = policy_cost(n_cr)
    policy["loop"] = policy_loop_cost(n_cr)

    per_cr = dict(pins=42, motors=42, bodies=43, points=92, grips_total=6)
    F = flops_per_creature_step(w.S, w.I, per_cr["pins"], per_cr["motors"], per_cr["bodies"], per_cr["points"],
                                g_att, r_near, per_cr["grips_total"])
    launch_ms = lat["device_mean_ms"]
    achieved_tflops = F * n_cr * chunk / (launch_ms * 1e-3) / 1e12
    timed_tflops = F * n_cr * args.steps / (dev_ms * 1e-3) / 1e12
    launch_bytes = algorithmic_launch_bytes(w, chunk)
    hbm_gbs = launch_bytes / (launch_ms * 1e-3) / 1e9
    # the kernel the launches ran: register-held round slots or slots read from LDS
    kernel = ("kx_k_persist_reg" if shape.get("register_rounds") else
              "kx_k_persist_fix" if shape.get("fixed_rounds") else "kx_k_persist")
    traffic = measured_traffic(n_cr, chunk, "uniform4096", kernel)
    plain = None
    if not args.no_plain:
        wp = W.world_uniform(args.creatures, seed=1)
        gp = G.create(wp, path=K.KX_PATH_PLAIN)
        G.upload(gp)
        G.rollout(gp, law, 0, 2)
        G.sync(gp)
        t1 = time.perf_counter()
        G.rollout(gp, law, 2, args.plain_steps)
        G.sync(gp)
        pwall = time.perf_counter() - t1
        pms, launches = G.last_timing(gp)
        plain = {"value": n_cr * args.plain_steps / pwall, "steps": args.plain_steps,
                 "ms_per_step": pwall * 1e3 / args.plain_steps, "device_ms_per_step": pms / args.plain_steps,
                 "launches_per_step": launches / args.plain_steps, "sdf": "linear scan (WGSL shape)"}
        G.set_plain_sdf(gp, True)
        t1 = time.perf_counter()
        G.rollout(gp, law, 2 + args.plain_steps, args.plain_steps)
        G.sync(gp)
        plain["value_grid_sdf"] = n_cr * args.plain_steps / (time.perf_counter() - t1)
        G.destroy(gp)

    out = {
        "metric": "creature-steps/sec on 1 MI355X (batched vs plain) at <=1e-3 px; p99 rollout ms",
        "value": value,
        "unit": "creature-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"uniform batch: {n_cr} creatures (6 arms x 7 segments), makeMap 1200x800 seed 1, "
                               f"S=4 x I=8, h=1/240, seeded control law + adhesion reflex",
                   "creatures": n_cr, "sim_steps": args.steps, "steps_per_launch": chunk,
                   "path": "persistent (one wavefront per creature, LDS-resident)"},
        "roofline": {"bound": "valu", "achieved": achieved_tflops, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved_tflops / FP32_PEAK_TFLOPS,
                     "traffic": traffic["traffic_bytes"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "algorithmic_bytes_per_launch": launch_bytes,
                     "flops_per_creature_step": F, "steps_per_launch": chunk, "creatures_per_launch": n_cr,
                     "grips_attached_mean": g_att, "rects_within_16px_mean": r_near,
                     "kernel": kernel, "launch_ms": launch_ms,
                     "launch_ms_source": f"mean of {lat['samples']} launches of {chunk} steps (HIP events)",
                     "achieved_timed_region": timed_tflops,
                     "hbm_algorithmic_gbs": hbm_gbs, "hbm_peak_gbs": HBM_PEAK_GBS,
                     "peak_unpacked": FP32_UNPACKED_TFLOPS, "frac_unpacked": achieved_tflops / FP32_UNPACKED_TFLOPS},
        "launch_shape": dict(shape, lanes_per_creature=args.lanes or None, waves_per_block=args.wpb or None),
        "p50_rollout_ms": lat["host_p50_ms"],
        "p99_rollout_ms": lat["host_p99_ms"],
        "rollout_unit": f"{chunk} steps x {n_cr} creatures per launch, submission to completion, "
                        f"{lat['samples']} launches",
        "rollout_latency": lat,
        "device_ms_total": dev_ms,
        "final_stats": st,
        "state_readout": readout,
        "policy": policy,
    }
    if plain:
        out["plain"] = plain
        out["speedup_vs_plain"] = value / plain["value"]
        out["speedup_vs_plain_grid_sdf"] = value / plain["value_grid_sdf"]
    if args.serve_requests:
        out["serving"] = serving_fixed_rate(args.serve_requests, args.serve_rate)
    if not args.no_configs:
        # bounded runs of the uneven and stress workloads (full runs:
        # --workload uneven / stress)
        out["uneven"] = uneven_leg()
        out["stress"] = stress_leg()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(device=(g, w, law, step))
        out["parity"] = out["cpu_baseline"]["device_parity"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
