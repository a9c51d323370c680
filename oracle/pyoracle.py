"""ctypes binding of the CPU oracle (oracle/*.c) — TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
always as the checker; the product package never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")

PIN, ANGLE, GRIP = 0, 1, 2
ARM_REACH = 24 + 7 * 13.5
MAX_ARMS = 8


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, i, u32 = C.c_double, C.c_int, C.c_uint32
        P = C.c_void_p
        dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
        sig = {
            "or_mb32": (u32, [u32]),
            "or_mulberry32_stream": (None, [u32, i, dp]),
            "or_ctl_u": (d, [u32, u32, u32, u32]),
            "or_js_hypot": (d, [d, d]),
            "or_make_map": (i, [d, d, i, i, d, d, u32, dp, i]),
            "or_sdf": (i, [dp, i, d, d, dp]),
            "or_world_new": (P, [i, i, d]),
            "or_world_free": (None, [P]),
            "or_world_set_map": (None, [P, d, d, dp, i]),
            "or_world_set_limits": (None, [P, d, d, d, d]),
            "or_spawn_kraken": (i, [P, d, d, d, i, i, i, d]),
            "or_find_spawn": (i, [P, d, C.POINTER(u32), dp]),
            "or_rng_next": (d, [C.POINTER(u32)]),
            "or_crn_spawns": (i, [P, i, i, i, C.POINTER(u32), dp]),
            "or_respawn": (None, [P, i, d, d, d]),
            "or_ensure_colored": (None, [P]),
            "or_permute_cons": (None, [P, ip]),
            "or_counts": (i, [P, ip]),
            "or_color_info": (None, [P, C.POINTER(u32), C.POINTER(u32)]),
            "or_get_bodies": (None, [P, dp, dp, dp, dp, dp, dp, dp, dp]),
            "or_set_bodies": (None, [P, dp, dp, dp, dp, dp, dp]),
            "or_get_pred": (None, [P, dp, dp, dp]),
            "or_set_pred": (None, [P, dp, dp, dp]),
            "or_get_cons": (None, [P, ip, dp]),
            "or_set_cons_dynamic": (None, [P, ip, dp, dp, dp, dp, dp]),
            "or_get_points": (None, [P, ip, dp]),
            "or_creature_info": (i, [P, i, ip]),
            "or_arm_tip": (i, [P, i, i]),
            "or_arm_grip_pos": (i, [P, i, i]),
            "or_get_cd": (None, [P, dp]),
            "or_set_cd": (None, [P, dp]),
            "or_attach_grip": (None, [P, i, i, d]),
            "or_apply_control": (None, [P, i, d, dp, dp, dp]),
            "or_apply_law": (None, [P, u32, u32]),
            "or_apply_law_one": (None, [P, u32, u32, i, u32]),
            "or_apply_law_ex": (None, [P, u32, u32, i, u32, C.POINTER(C.c_double)]),
            "or_step": (None, [P]),
            "or_substep": (None, [P]),
            "or_solve_collisions": (None, [P]),
            "or_get_stats": (None, [P, dp]),
            "or_f32ify": (None, [P]),
            "or_events": (i, [P, ip, i]),
            "or_clear_events": (None, [P]),
            "or_set_step_counter": (None, [P, i]),
            "or_bench_threads": (d, [P, i, i, u32]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def mulberry32(seed, n):
    out = np.zeros(n)
    lib().or_mulberry32_stream(seed & 0xFFFFFFFF, n, _dp(out))
    return out


def ctl_u(seed, creature, step, slot):
    return lib().or_ctl_u(seed & 0xFFFFFFFF, creature, step, slot)


def make_map(width, height, seed=1, cols=4, rows=3, density=0.48, wall=14.0):
    cap = 4 * (cols + 1) * (rows + 1) * 2 + 16
    buf = np.zeros(cap * 4)
    n = lib().or_make_map(width, height, cols, rows, density, wall, seed & 0xFFFFFFFF, _dp(buf), cap)
    assert n <= cap
    return buf[: 4 * n].reshape(n, 4).copy()


def sdf(rects, x, y):
    rects = np.ascontiguousarray(rects, dtype=np.float64).reshape(-1, 4)
    out = np.zeros(3)
    ok = lib().or_sdf(_dp(rects), len(rects), x, y, _dp(out))
    return tuple(out) if ok else None


class World:
    """One oracle world (f64 by default, f32 emulation after f32ify())."""

    def __init__(self, substeps=4, iterations=8, h=1 / 240):
        self.p = lib().or_world_new(substeps, iterations, h)
        self.S, self.I, self.h = substeps, iterations, h

    def __del__(self):
        if getattr(self, "p", None):
            lib().or_world_free(self.p)
            self.p = None

    def set_map(self, width, height, rects):
        rects = np.ascontiguousarray(rects, dtype=np.float64).reshape(-1, 4)
        self.rects = rects
        lib().or_world_set_map(self.p, width, height, _dp(rects), len(rects))

    def spawn(self, x, y, angle, arms=6, segs=7, ring=8, curl=0.28):
        r = lib().or_spawn_kraken(self.p, x, y, angle, arms, segs, ring, curl)
        assert r >= 0
        return r

    def find_spawn(self, clearance, rng_state):
        st = C.c_uint32(rng_state[0])
        xy = np.zeros(2)
        lib().or_find_spawn(self.p, clearance, C.byref(st), _dp(xy))
        rng_state[0] = st.value
        return xy

    def crn_spawns(self, pop, rng_state, pair_shared=True, has_map=True):
        """crnSpawns (src/sim/train.js:103-122) from the stream state
        rng_state[0] (advanced in place): [pop, 3] x, y, angle, and the
        number of stream calls made."""
        st = C.c_uint32(rng_state[0])
        out = np.zeros(3 * max(pop, 1))
        calls = lib().or_crn_spawns(self.p, 1 if has_map else 0, pop, 1 if pair_shared else 0, C.byref(st), _dp(out))
        rng_state[0] = st.value
        return out[: 3 * pop].reshape(pop, 3), calls

    def respawn(self, ci, x, y, angle):
        """respawnCreature (src/sim/train.js:188-229) of creature ci."""
        lib().or_respawn(self.p, ci, x, y, angle)

    def color(self):
        lib().or_ensure_colored(self.p)

    def permute_cons(self, order):
        o = np.ascontiguousarray(order, dtype=np.int32)
        lib().or_permute_cons(self.p, _ip(o))

    def counts(self):
        o = np.zeros(4, dtype=np.int32)
        lib().or_counts(self.p, _ip(o))
        return tuple(int(v) for v in o)

    @property
    def n(self):
        return self.counts()[0]

    def color_info(self):
        off = (C.c_uint32 * 16)()
        cnt = (C.c_uint32 * 16)()
        lib().or_color_info(self.p, off, cnt)
        return list(off), list(cnt)

    def bodies(self):
        n = self.n
        a = [np.zeros(n) for _ in range(8)]
        lib().or_get_bodies(self.p, *[_dp(x) for x in a])
        return dict(zip(["px", "py", "ang", "vx", "vy", "av", "im", "ii"], a))

    def set_bodies(self, px, py, ang, vx, vy, av):
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (px, py, ang, vx, vy, av)]
        lib().or_set_bodies(self.p, *[_dp(x) for x in arrs])

    def pred(self):
        n = self.n
        a = [np.zeros(n) for _ in range(3)]
        lib().or_get_pred(self.p, *[_dp(x) for x in a])
        return dict(zip(["ppx", "ppy", "angP"], a))

    def set_pred(self, ppx, ppy, angP):
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (ppx, ppy, angP)]
        lib().or_set_pred(self.p, *[_dp(x) for x in arrs])

    def cons(self):
        nc = self.counts()[1]
        I = np.zeros(nc * 6, dtype=np.int32)
        D = np.zeros(nc * 15)
        lib().or_get_cons(self.p, _ip(I), _dp(D))
        I = I.reshape(nc, 6)
        D = D.reshape(nc, 15)
        names_i = ["type", "a", "b", "owner", "color", "broken"]
        names_d = ["lax", "lay", "lbx", "lby", "compliance", "rest", "target", "minA", "maxA",
                   "lambda", "lambdaX", "lambdaY", "breakForce", "force", "id"]
        out = {k: I[:, j].copy() for j, k in enumerate(names_i)}
        out.update({k: D[:, j].copy() for j, k in enumerate(names_d)})
        out["id"] = out["id"].astype(np.int64)
        return out

    def set_cons_dynamic(self, broken=None, lbx=None, lby=None, compliance=None, target=None, breakForce=None):
        def dp(v):
            return None if v is None else _dp(np.ascontiguousarray(v, dtype=np.float64))
        keep = [np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
                for v in (lbx, lby, compliance, target, breakForce)]
        b = None if broken is None else np.ascontiguousarray(broken, dtype=np.int32)
        lib().or_set_cons_dynamic(self.p, None if b is None else _ip(b),
                                  *[None if v is None else _dp(v) for v in keep])

    def points(self):
        np_ = self.counts()[2]
        b = np.zeros(np_, dtype=np.int32)
        d = np.zeros(np_ * 3)
        lib().or_get_points(self.p, _ip(b), _dp(d))
        d = d.reshape(np_, 3)
        return {"body": b, "lx": d[:, 0].copy(), "ly": d[:, 1].copy(), "r": d[:, 2].copy()}

    def creature(self, ci):
        o = np.zeros(6, dtype=np.int32)
        lib().or_creature_info(self.p, ci, _ip(o))
        return dict(zip(["mantle", "arms", "segs", "ring", "first_body", "nbodies"], (int(v) for v in o)))

    def arm_tip(self, ci, a):
        return lib().or_arm_tip(self.p, ci, a)

    def grip_pos(self, ci, a):
        return lib().or_arm_grip_pos(self.p, ci, a)

    def cd(self):
        ncr = self.counts()[3]
        out = np.zeros(ncr * MAX_ARMS)
        lib().or_get_cd(self.p, _dp(out))
        return out.reshape(ncr, MAX_ARMS)

    def set_cd(self, cd):
        cd = np.ascontiguousarray(cd, dtype=np.float64)
        lib().or_set_cd(self.p, _dp(cd))

    def attach_grip(self, ci, a, bf):
        lib().or_attach_grip(self.p, ci, a, bf)

    def apply_control(self, ci, tone, shoulder, curl, adhesion):
        s = np.ascontiguousarray(shoulder, dtype=np.float64)
        c = np.ascontiguousarray(curl, dtype=np.float64)
        a = np.ascontiguousarray(adhesion, dtype=np.float64)
        lib().or_apply_control(self.p, ci, tone, _dp(s), _dp(c), _dp(a))

    def apply_law(self, seed, step):
        lib().or_apply_law(self.p, seed & 0xFFFFFFFF, step)

    def apply_law_ex(self, seed, step, creature, key, grip_min, grip_max, attach_min, release_below, fixed_grip):
        """The control law with a test grip range / thresholds for one creature."""
        p = (C.c_double * 5)(grip_min, grip_max, attach_min, release_below, 1.0 if fixed_grip else 0.0)
        lib().or_apply_law_ex(self.p, seed & 0xFFFFFFFF, step, creature, key & 0xFFFFFFFF, p)

    def step(self):
        lib().or_step(self.p)

    def substep(self):
        lib().or_substep(self.p)

    def solve_collisions(self):
        lib().or_solve_collisions(self.p)

    def stats(self):
        o = np.zeros(5)
        lib().or_get_stats(self.p, _dp(o))
        return dict(zip(["maxSpeed", "maxAngSpeed", "nan", "collisions", "broken"], o))

    def f32ify(self):
        lib().or_f32ify(self.p)

    def events(self):
        n = lib().or_events(self.p, None, 0)
        buf = np.zeros(max(n, 1) * 4, dtype=np.int32)
        lib().or_events(self.p, _ip(buf), n)
        return [tuple(int(v) for v in buf[4 * i: 4 * i + 4]) for i in range(n)]

    def clear_events(self):
        lib().or_clear_events(self.p)

    def set_step_counter(self, s):
        lib().or_set_step_counter(self.p, s)

    def bench_threads(self, threads, steps, seed=1):
        return lib().or_bench_threads(self.p, threads, steps, seed & 0xFFFFFFFF)


# ── proprioception readout (restated from the reference; checker only) ──

SEG_LEN = 13.5


def read_proprio(px, py, ang, vx, vy, av, mantle, arms, grips, prev, h, substeps):
    """readProprio (src/sim/core.js:129-175) of one creature, in f64.

    arms: per arm its segment bodies (segs[0] .. tip); grips: per arm
    (broken, force, breakForce or None for unbreakable); prev: the previous
    call's mantle velocity (prevVx, prevVy). Returns (out, side) with out the
    6*A + 6 inputs and side = (meanExtension, meanCurl, anchoredCount, speed,
    (vx, vy) to carry as the next prev). Helpers: tipPosition, armCurl and
    armExtension (src/body/kraken.js:150-183), readState (core.js:100-108).
    """
    import math
    shoulder_limit, joint_limit = math.pi * 0.62, math.pi * 0.42
    x, y = float(px[mantle]), float(py[mantle])
    cvx, cvy = float(vx[mantle]), float(vy[mantle])
    angle, cav = float(ang[mantle]), float(av[mantle])
    speed = math.hypot(cvx, cvy)
    ca, sa = math.cos(-angle), math.sin(-angle)

    def tip(segs):
        b = segs[-1]
        L = SEG_LEN * 0.5
        return float(px[b]) + L * math.cos(float(ang[b])), float(py[b]) + L * math.sin(float(ang[b]))

    def curl(segs):
        s, n = 0.0, 0
        for i in range(1, len(segs)):
            d = float(ang[segs[i]]) - float(ang[segs[i - 1]])
            s += math.atan2(math.sin(d), math.cos(d))
            n += 1
        return s / n if n else 0.0

    def extension(segs):
        tx, ty = tip(segs)
        s0 = segs[0]
        return min(1.0, math.hypot(tx - float(px[s0]), ty - float(py[s0])) / max(1.0, ARM_REACH))

    def cl(v, m):
        return m if v > m else (-m if v < -m else v)

    def jsmin(a, b):
        return math.nan if (a != a or b != b) else (a if a < b else b)

    def jsmax(a, b):
        return math.nan if (a != a or b != b) else (a if a > b else b)

    A = len(arms)
    anchored, ext_sum, curl_sum = 0, 0.0, 0.0
    shoulder = []
    for a in range(A):
        d = float(ang[arms[a][0]]) - angle
        shoulder.append(math.atan2(math.sin(d), math.cos(d)))
        ext_sum += extension(arms[a])
        curl_sum += curl(arms[a])
        if not grips[a][0]:
            anchored += 1
    out = []
    out += [cl(shoulder[a] / shoulder_limit, 1.2) for a in range(A)]
    out += [cl(curl(arms[a]) / joint_limit, 1.2) for a in range(A)]
    for a in range(A):
        tx, ty = tip(arms[a])
        dx, dy = tx - x, ty - y
        out.append(cl((dx * ca - dy * sa) / ARM_REACH, 1.2))
        out.append(cl((dx * sa + dy * ca) / ARM_REACH, 1.2))
    for a in range(A):
        broken, force, bf = grips[a]
        f = force if (force == force and force) else 0.0
        bfv = math.inf if bf is None else bf
        out.append(0.0 if broken else jsmin(1.0, f / jsmax(1.0, bfv)))
    out += [0.0 if grips[a][0] else 1.0 for a in range(A)]
    ax = (cvx - prev[0]) / (h * substeps)
    ay = (cvy - prev[1]) / (h * substeps)
    out.append(jsmax(-1.0, jsmin(1.0, (ax * ca - ay * sa) / 600)))
    out.append(jsmax(-1.0, jsmin(1.0, (ax * sa + ay * ca) / 600)))
    out.append(anchored / A)
    mean_ext, mean_curl = ext_sum / A, curl_sum / A
    out.append(mean_ext)
    out.append(jsmin(1.0, speed / 300))
    out.append(jsmax(-1.0, jsmin(1.0, cav / 18)))
    return out, (mean_ext, mean_curl, anchored, speed, (cvx, cvy))


# ── policy network and ES (src/brain/net.js, src/brain/es.js), numpy f64 ──

HIDDEN = 24                      # src/brain/net.js:20
OUT_DIM = 1 + 6 * 3 + 1          # src/brain/net.js:23


def weight_count(n_in, hidden=HIDDEN, n_out=OUT_DIM):
    """src/brain/net.js:30-32."""
    return n_in * hidden + hidden + hidden * hidden + hidden + hidden * n_out + n_out


def _act(s):
    # net.js:81/87/93: s > 20 -> 1, s < -20 -> -1, else tanh(s); NaN stays NaN
    return np.where(s > 20, 1.0, np.where(s < -20, -1.0, np.tanh(s)))


def forward_into(w, x, n_in, hidden=HIDDEN, n_out=OUT_DIM):
    """forwardInto (src/brain/net.js:71-98) for every creature of a population:
    w float32 [pop, nW] (per-creature genomes), x float32 [pop, n_in] ->
    (out float32 [pop, n_out], scratch float32 [pop, 2*hidden]). Each sum
    starts at its bias and adds x[i]*w[row+i] for i = 0, 1, ... in float64
    (products of two float32 values are exact in float64), the activations are
    stored as float32 as the reference's Float32Array scratch and outputs are.
    The layer-2 bias is read at offset b2 + j, the first row of the layer-2
    weights, as net.js:84-86 reads it (the stored layer-2 biases at b3 are
    never used)."""
    w = np.asarray(w, np.float32).astype(np.float64)
    x = np.asarray(x, np.float32).astype(np.float64)
    pop = w.shape[0]
    b1 = n_in * hidden
    b2 = b1 + hidden
    b3 = b2 + hidden * hidden
    b4 = b3 + hidden
    b5 = b4 + hidden * n_out
    w1 = w[:, :b1].reshape(pop, hidden, n_in)
    s = w[:, b1:b1 + hidden].copy()
    for i in range(n_in):
        s = s + x[:, None, i] * w1[:, :, i]
    h1 = _act(s).astype(np.float32).astype(np.float64)
    w2 = w[:, b2:b2 + hidden * hidden].reshape(pop, hidden, hidden)
    s = w[:, b2:b2 + hidden].copy()
    for i in range(hidden):
        s = s + h1[:, None, i] * w2[:, :, i]
    h2 = _act(s).astype(np.float32).astype(np.float64)
    w3 = w[:, b4:b5].reshape(pop, n_out, hidden)
    s = w[:, b5:b5 + n_out].copy()
    for i in range(hidden):
        s = s + h2[:, None, i] * w3[:, :, i]
    out = _act(s).astype(np.float32)
    return out, np.concatenate([h1, h2], axis=1).astype(np.float32)


def decode(out, arms):
    """decode (src/brain/net.js:114-124) as a row per creature:
    [tone, (shoulder, curl, adhesion) per arm, emitEcho] in float64."""
    out = np.asarray(out, np.float32).astype(np.float64)
    rec = np.zeros((out.shape[0], 2 + 3 * arms))
    rec[:, 0] = 0.5 + 0.5 * out[:, 0]
    for a in range(arms):
        o = 1 + a * 3
        rec[:, 1 + 3 * a] = out[:, o]
        rec[:, 2 + 3 * a] = out[:, o + 1]
        rec[:, 3 + 3 * a] = np.where(out[:, o + 2] > 0, out[:, o + 2], 0.0)
    rec[:, -1] = out[:, OUT_DIM - 1] > 0.3
    return rec


def mulberry32_at(state, idx):
    """The value of call idx (0-based) of a mulberry32 stream whose state is
    `state` before the first call (src/world/map.js:11-20; the trainer's stream
    of src/sim/train.js:65-72 is the same generator): the state advances by
    0x6D2B79F5 per call, so any call's value follows from its index."""
    a = (np.uint64(state) + (np.asarray(idx, np.uint64) + np.uint64(1)) * np.uint64(0x6D2B79F5)) \
        & np.uint64(0xFFFFFFFF)
    t = a.astype(np.uint32)
    with np.errstate(over="ignore"):
        t = (t ^ (t >> np.uint32(15))) * (t | np.uint32(1))
        t = t ^ (t + (t ^ (t >> np.uint32(7))) * (t | np.uint32(61)))
        t = t ^ (t >> np.uint32(14))
    return t.astype(np.float64) / 4294967296.0


def genome_generation(theta, sigma, pop, state):
    """beginGeneration's genomeAt loop (src/sim/train.js:154-158,
    src/brain/es.js:67-88) for i = 0 .. pop-1 from a trainer stream whose
    state is `state`: pair p = i >> 1 draws eps_p (Irwin-Hall, four calls per
    weight, k-major) on its first member; genome i = theta + eps_p * sigma *
    (+1 for even i, -1 for odd), rounded to float32. Returns (genomes float32
    [pop, nW], eps_bank float32 [pop//2, nW], state after the draws)."""
    theta = np.asarray(theta, np.float32)
    nw = theta.shape[0]
    pairs = pop >> 1
    base = np.arange(pairs * nw, dtype=np.uint64) * np.uint64(4)
    r = [mulberry32_at(state, base + np.uint64(d)) for d in range(4)]
    eps = ((((r[0] + r[1]) + r[2]) + r[3]) - 2) * 1.7320508
    eps = eps.astype(np.float32).reshape(pairs, nw)
    gen = np.zeros((pop, nw), np.float32)
    th = theta.astype(np.float64)
    for i in range(pop):
        s = sigma * (-1.0 if i & 1 else 1.0)
        gen[i] = (th + eps[i >> 1].astype(np.float64) * s).astype(np.float32)
    state2 = (int(state) + pairs * nw * 4 * 0x6D2B79F5) & 0xFFFFFFFF
    return gen, eps, state2


def es_update(theta, m, v, t, eps_bank, fitness, opts):
    """update (src/brain/es.js:92-149): mirrored-pair advantages normalised
    over the generation, the clipped gradient accumulated in float32 pair by
    pair, one Adam step. theta, m, v float32 [nW] (copies are returned), t the
    step count before the update, opts the es.js option names. Returns
    (theta, m, v, t, sigma, stats{best, mean, sd, sigma})."""
    fitness = np.asarray(fitness, np.float64)
    pop = fitness.shape[0]
    half = pop >> 1
    diffs = (fitness[0::2][:half] - fitness[1::2][:half]) / 2
    mean = 0.0
    for p in range(half):
        mean += fitness[p * 2] + fitness[p * 2 + 1]
    mean /= opts["popSize"]
    sd = 0.0
    for p in range(half):
        sd += diffs[p] * diffs[p]
    sd = float(np.sqrt(sd / max(1, half))) or 1.0
    inv = 1 / (half * opts["sigma"] * sd)
    g = np.zeros(theta.shape[0], np.float32)
    eb = np.asarray(eps_bank, np.float32).astype(np.float64)
    for p in range(half):
        g = (g.astype(np.float64) + diffs[p] * eb[p]).astype(np.float32)
    a = g.astype(np.float64) * inv
    a = np.where(a > opts["clip"], opts["clip"], np.where(a < -opts["clip"], -opts["clip"], a))
    g = a.astype(np.float32).astype(np.float64)
    t = t + 1
    b1, b2 = opts["adamBeta1"], opts["adamBeta2"]
    c1, c2 = 1 - b1 ** t, 1 - b2 ** t
    m = (b1 * np.asarray(m, np.float32).astype(np.float64) + (1 - b1) * g).astype(np.float32)
    v = (b2 * np.asarray(v, np.float32).astype(np.float64) + (1 - b2) * g * g).astype(np.float32)
    with np.errstate(all="ignore"):
        step = opts["lr"] * (m.astype(np.float64) / c1) / (np.sqrt(v.astype(np.float64) / c2) + opts["adamEps"])
    step = np.where(np.isfinite(step), step, 0.0)
    theta = (np.asarray(theta, np.float32).astype(np.float64) + step).astype(np.float32)
    sigma = max(opts["minSigma"], opts["sigma"] * opts["sigmaDecay"])
    best = float(np.max(fitness))
    return theta, m, v, t, sigma, dict(best=best, mean=mean, sd=sd * 2, sigma=sigma)
