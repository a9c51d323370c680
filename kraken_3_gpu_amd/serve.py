"""Serving loop over the native continuous-batching server (kx_server_*).

`Server` owns a kx_server and keeps each request's output arrays alive until
its completion has been polled. `es_stream` draws the ES request stream:
Poisson arrivals at a fixed rate, 16-512 antithetic pairs, 300-1500 steps, a
morphology from the uneven range (4-8 arms x 5-9 segments, mantle ring 6-10)
and a spawn pose from findSpawn on the shared map. `run_stream` submits a
stream at its arrival times from a feeder thread, polls the completions and
returns latency percentiles beside the aggregate creature-steps/s.
"""
import ctypes as C
import math
import threading
import time

import numpy as np

from . import _lib as K
from . import world as W

MAP = dict(width=1200.0, height=800.0, cols=4, rows=3, density=0.48, wall=14.0, seed=1)
SIGMA = 0.06          # es.js default perturbation scale


def default_opts(chunk_steps=50, max_active=64, host_threads=0, hbm_reserve=2 << 30, law=None, map_=None):
    o = K.ServerOpts()
    o.chunk_steps, o.max_active, o.host_threads, o.hbm_reserve = chunk_steps, max_active, host_threads, hbm_reserve
    s = W.SOLVER
    p = o.params
    p.substeps, p.iterations, p.h = s["substeps"], s["iterations"], s["h"]
    p.lin_drag, p.ang_drag, p.max_speed, p.max_ang_speed = (s["lin_drag"], s["ang_drag"], s["max_speed"],
                                                            s["max_ang_speed"])
    p.path, p.lanes_per_creature, p.waves_per_block = K.KX_PATH_PERSISTENT, 16, 1
    o.law = law or W.default_law(0)
    m = dict(MAP, **(map_ or {}))
    o.map_width, o.map_height, o.map_density, o.map_wall = m["width"], m["height"], m["density"], m["wall"]
    o.map_cols, o.map_rows, o.map_seed = m["cols"], m["rows"], m["seed"]
    return o


def bodies_per_creature(arms, segs):
    """spawnKraken's body count: the mantle plus one body per arm segment."""
    return 1 + arms * segs


def make_request(d):
    """kx_request from a dict (arms, segs, ring, pairs, steps, x, y, angle,
    policy_seed, noise_seed, sigma, tag)."""
    r = K.Request()
    r.arms, r.segs, r.ring, r.pairs, r.steps = d["arms"], d["segs"], d["ring"], d["pairs"], d["steps"]
    r.x, r.y, r.angle = d["x"], d["y"], d["angle"]
    r.policy_seed, r.noise_seed = d["policy_seed"] & 0xFFFFFFFF, d["noise_seed"] & 0xFFFFFFFF
    r.sigma, r.tag = d.get("sigma", SIGMA), d.get("tag", 0)
    return r


def policy_tape(d, step0, nsteps, step_seconds=None):
    """The control-tape records the server feeds request d for steps
    step0 .. step0+nsteps-1: float32 [nsteps, 2*pairs, KX_TAPE_WORDS]."""
    if step_seconds is None:
        step_seconds = W.SOLVER["h"] * W.SOLVER["substeps"]
    r = make_request(d)
    t = np.zeros((nsteps, 2 * d["pairs"], K.KX_TAPE_WORDS), np.float32)
    K.check(K.lib().kx_policy_tape(C.byref(r), step0, nsteps, step_seconds, K.ptr(t, C.c_float)))
    return t


class Server:
    def __init__(self, opts=None):
        self.opts = opts or default_opts()
        self.p = C.c_void_p()
        K.check(K.lib().kx_server_create(C.byref(self.opts), C.byref(self.p)))
        self._keep = {}
        self._tag = 0

    def submit(self, d, want_state=False):
        """Queues request dict d; returns its tag."""
        self._tag += 1
        tag = d.get("tag", self._tag)
        d = dict(d, tag=tag)
        r = make_request(d)
        ncr = 2 * d["pairs"]
        fit = np.zeros(ncr, np.float32)
        r.fitness = K.ptr(fit, C.c_float)
        state = None
        if want_state:
            nb = ncr * bodies_per_creature(d["arms"], d["segs"])
            state = np.zeros(6 * nb, np.float32)
            r.final_state, r.final_cap = K.ptr(state, C.c_float), state.size
        self._keep[tag] = (r, fit, state, d)
        K.check(K.lib().kx_server_submit(self.p, C.byref(r)))
        return tag

    def poll(self, wait=0.0, cap=64):
        """Completed requests: dicts with tag, status, n_creatures, n_bodies,
        t_submit/t_admit/t_done (server clock, s), device_bytes, fitness and
        (when asked for) final_state {px, py, ang, vx, vy, av}."""
        buf = (K.RequestDone * cap)()
        n = C.c_int32()
        K.check(K.lib().kx_server_poll(self.p, buf, cap, C.byref(n), wait))
        out = []
        for i in range(n.value):
            e = buf[i]
            _, fit, state, d = self._keep.pop(e.tag)
            res = dict(tag=e.tag, status=e.status, n_creatures=e.n_creatures, n_bodies=e.n_bodies,
                       t_submit=e.t_submit, t_admit=e.t_admit, t_done=e.t_done, device_bytes=e.device_bytes,
                       fitness=fit, request=d)
            if state is not None and e.status == 0:
                nb = e.n_bodies
                res["final_state"] = {k: state[j * nb:(j + 1) * nb] for j, k in
                                      enumerate(("px", "py", "ang", "vx", "vy", "av"))}
            out.append(res)
        return out

    def info(self):
        i = K.ServerInfo()
        K.check(K.lib().kx_server_info_get(self.p, C.byref(i)))
        return {k: getattr(i, k) for k, _ in K.ServerInfo._fields_}

    def close(self):
        if self.p:
            K.lib().kx_server_destroy(self.p)
            self.p = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def es_stream(n, rate, seed=1, pairs=(16, 512), steps=(300, 1500), map_=None):
    """n ES requests with Poisson arrivals at `rate` per second: arrival
    offsets (s) and request dicts. Morphology, sizes, pose and seeds are all
    drawn from `seed`."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n) if rate > 0 else np.zeros(n)
    t = np.cumsum(gaps) - gaps[0]
    m = dict(MAP, **(map_ or {}))
    b = W.Builder()
    b.map(m["width"], m["height"], m["seed"], m["cols"], m["rows"], m["density"], m["wall"])
    st = [(seed ^ 0x9E3779B9) & 0xFFFFFFFF]
    reqs = []
    for i in range(n):
        arms, segs, ring = W.morphology(seed, i)
        xy = b.find_spawn(W.ARM_REACH + 20, st)
        reqs.append(dict(arms=arms, segs=segs, ring=ring, pairs=int(rng.integers(pairs[0], pairs[1] + 1)),
                         steps=int(rng.integers(steps[0], steps[1] + 1)), x=float(xy[0]), y=float(xy[1]),
                         angle=float(rng.random() * 2 * math.pi), policy_seed=seed * 7919 + i,
                         noise_seed=seed * 104729 + 31 * i, sigma=SIGMA, tag=i + 1))
    return t, reqs


def saturate(server, reqs, outstanding=48, warm_s=2.0, measure_s=4.0):
    """Serving capacity with the queue never empty: keeps `outstanding`
    requests submitted (a new one as each completes, cycling through reqs)
    and returns the creature-steps/s the server executed over `measure_s`
    seconds after `warm_s` seconds of warm-up, with the device-busy fraction
    of that window."""
    it = iter(reqs * 1000)
    tag = [0]
    live = [0]

    def top_up():
        while live[0] < outstanding:
            tag[0] += 1
            server.submit(dict(next(it), tag=1_000_000 + tag[0]))
            live[0] += 1
    top_up()
    t0 = time.perf_counter()
    marks = []
    while True:
        live[0] -= len(server.poll(wait=0.02))
        top_up()
        el = time.perf_counter() - t0
        if not marks and el >= warm_s:
            marks.append(server.info())
        if marks and el >= warm_s + measure_s:
            marks.append(server.info())
            break
    a, b = marks
    dt = b["now"] - a["now"]
    return dict(creature_steps_per_s=(b["creature_steps"] - a["creature_steps"]) / dt,
                device_busy_frac=(b["busy"] - a["busy"]) / dt, window_s=dt, outstanding=outstanding,
                launches=b["launches"] - a["launches"])


def run_stream(server, arrivals, reqs, want_state=False, timeout=600.0):
    """Submits reqs at their arrival offsets (relative to now) and collects
    every completion. Returns (results sorted by tag, summary)."""
    start = time.perf_counter()
    err = []

    def feed():
        try:
            for t, d in zip(arrivals, reqs):
                delay = start + t - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                server.submit(d, want_state)
        except Exception as e:       # reported by the caller
            err.append(e)

    th = threading.Thread(target=feed)
    th.start()
    results = []
    while len(results) < len(reqs) and time.perf_counter() - start < timeout and not err:
        results.extend(server.poll(wait=0.05))
    th.join()
    if err:
        raise err[0]
    if len(results) < len(reqs):
        raise TimeoutError(f"{len(results)} of {len(reqs)} requests completed in {timeout} s")
    results.sort(key=lambda r: r["tag"])
    lat = np.array([r["t_done"] - r["t_submit"] for r in results])
    wait = np.array([r["t_admit"] - r["t_submit"] for r in results])
    cs = sum(r["request"]["steps"] * r["n_creatures"] for r in results)
    span = max(r["t_done"] for r in results) - min(r["t_submit"] for r in results)
    info = server.info()
    summary = dict(requests=len(results), failed=sum(r["status"] != 0 for r in results),
                   creature_steps=int(cs), span_s=span, creature_steps_per_s=cs / span if span > 0 else 0.0,
                   p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                   max_ms=float(lat.max() * 1e3), p50_admit_wait_ms=float(np.percentile(wait, 50) * 1e3),
                   p99_admit_wait_ms=float(np.percentile(wait, 99) * 1e3),
                   device_busy_frac=info["busy"] / span if span > 0 else 0.0, launches=info["launches"],
                   peak_active=info["peak_active"], admit_waits=info["admit_waits"])
    return results, summary
