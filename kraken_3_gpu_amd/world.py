"""Caller-side world state and synthetic workloads.

`World` is the Python counterpart of the reference's XpbdWorld as the GPU
boundary sees it (src/physics/xpbd.js:188-240): f64 body arrays, a flat
constraint table, collision points, the map's rect list, plus the creature
structure (arms, tips, grips, motors) that spawnKraken returns
(src/body/kraken.js:64-148). Worlds are generated by the library's C++
builder (kx_builder_*), which restates spawnKraken / makeMap / findSpawn.
"""
import ctypes as C
import math

import numpy as np

from . import _lib as K

ARM_REACH = 24 + 7 * 13.5          # kraken.js:49
SHOULDER_LIMIT = math.pi * 0.62     # kraken.js:35
JOINT_LIMIT = math.pi * 0.42        # kraken.js:34

# reference defaults (src/physics/xpbd.js:46-54) with the configs' S=4, I=8
SOLVER = dict(substeps=4, iterations=8, h=1 / 240, lin_drag=3.0, ang_drag=3.0, max_speed=360.0, max_ang_speed=18.0)


def mulberry32(seed):
    """Stateful mulberry32 (src/world/map.js:11-20)."""
    a = [seed & 0xFFFFFFFF]

    def rnd():
        a[0] = (a[0] + 0x6D2B79F5) & 0xFFFFFFFF
        t = a[0]
        t = ((t ^ (t >> 15)) * (t | 1)) & 0xFFFFFFFF
        t ^= (t + (((t ^ (t >> 7)) * (t | 61)) & 0xFFFFFFFF)) & 0xFFFFFFFF
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0
    rnd.state = a
    return rnd


class World:
    """f64 caller-side world. Constraint arrays are in list order; after
    gpuworld.create they are in colour order, like ensureColored leaves an
    XpbdWorld (xpbd.js:108-126)."""

    CONS_F = ("lax", "lay", "lbx", "lby", "compliance", "rest", "minA", "maxA", "target", "breakForce",
              "lambda_", "lambdaX", "lambdaY", "force")

    def __init__(self, builder, map_w, map_h, solver=None):
        L = K.lib()
        self.solver = dict(SOLVER, **(solver or {}))
        self.S = self.solver["substeps"]
        self.I = self.solver["iterations"]
        self.h = self.solver["h"]
        desc, cr = K.WorldDesc(), K.CreatureDesc()
        K.check(L.kx_builder_desc(builder, C.byref(desc), C.byref(cr)))
        nb, nc, npt, nr = desc.n_bodies, desc.n_constraints, desc.n_points, desc.n_rects
        self.n = nb
        self.width, self.height = map_w, map_h

        def cp(p, n, dt):
            return np.ctypeslib.as_array(p, shape=(n,)).astype(dt).copy() if n else np.zeros(0, dt)
        self.px, self.py, self.ang = (np.zeros(nb) for _ in range(3))
        K.check(L.kx_builder_state(builder, K.ptr(self.px, C.c_double), K.ptr(self.py, C.c_double),
                                   K.ptr(self.ang, C.c_double)))
        self.vx, self.vy, self.av = (np.zeros(nb) for _ in range(3))
        self.fx, self.fy, self.tq = (np.zeros(nb) for _ in range(3))
        self.im = cp(desc.inv_mass, nb, np.float32)
        self.ii = cp(desc.inv_inertia, nb, np.float32)
        self.type = cp(desc.c_type, nc, np.int32)
        self.a = cp(desc.c_a, nc, np.int32)
        self.b = cp(desc.c_b, nc, np.int32)
        self.owner = cp(desc.c_owner, nc, np.int32)
        self.lax = cp(desc.c_lax, nc, np.float32).astype(np.float64)
        self.lay = cp(desc.c_lay, nc, np.float32).astype(np.float64)
        self.rest = cp(desc.c_rest, nc, np.float32).astype(np.float64)
        self.minA = cp(desc.c_min, nc, np.float32).astype(np.float64)
        self.maxA = cp(desc.c_max, nc, np.float32).astype(np.float64)
        self.lbx, self.lby, self.compliance, self.target, self.breakForce = (np.zeros(nc) for _ in range(5))
        self.broken = np.zeros(nc, dtype=np.uint8)
        K.check(L.kx_builder_cons_init(builder, *(K.ptr(x, C.c_double) for x in
                                                  (self.lbx, self.lby, self.compliance, self.target, self.breakForce)),
                                       K.ptr(self.broken, C.c_uint8)))
        self.lambda_, self.lambdaX, self.lambdaY, self.force = (np.zeros(nc) for _ in range(4))
        self.color = np.full(nc, -1, dtype=np.int32)
        self.p_body = cp(desc.p_body, npt, np.int32)
        self.p_lx = cp(desc.p_lx, npt, np.float32)
        self.p_ly = cp(desc.p_ly, npt, np.float32)
        self.p_r = cp(desc.p_r, npt, np.float32)
        self.rects = cp(desc.rects, 4 * nr, np.float32).reshape(nr, 4)
        ncr = cr.n_creatures
        self.arm_off = cp(cr.creature_arm_off, ncr + 1, np.int32)
        self.arm_tip = cp(cr.arm_tip, cr.n_arms, np.int32)
        self.arm_grip = cp(cr.arm_grip, cr.n_arms, np.int32)
        self.motor_off = cp(cr.arm_motor_off, cr.n_arms + 1, np.int32)
        self.motors = cp(cr.motors, int(self.motor_off[-1]) if cr.n_arms else 0, np.int32)
        self.n_creatures = ncr
        self.cd = np.zeros((ncr, 8))
        self.colorOff = self.colorCnt = None
        self.stats = dict(maxSpeed=0.0, maxAngSpeed=0.0, nan=0)

    @property
    def nc(self):
        return len(self.type)

    def reorder_constraints(self, order):
        """Apply a permutation (new k -> old index) to every constraint array
        and to the creature tables that index constraints."""
        order = np.asarray(order)
        for name in ("type", "a", "b", "owner", "broken", "color") + self.CONS_F:
            setattr(self, name, getattr(self, name)[order].copy())
        inv = np.empty_like(order)
        inv[order] = np.arange(len(order))
        self.arm_grip = inv[self.arm_grip].astype(np.int32)
        self.motors = inv[self.motors].astype(np.int32)

    def desc(self, keep):
        """kx_world_desc over f32 copies (kept alive in `keep`)."""
        def f(x):
            a = np.ascontiguousarray(x, dtype=np.float32)
            keep.append(a)
            return K.ptr(a, C.c_float)

        def i(x):
            a = np.ascontiguousarray(x, dtype=np.int32)
            keep.append(a)
            return K.ptr(a, C.c_int32)
        d = K.WorldDesc()
        d.n_bodies, d.n_constraints, d.n_points, d.n_rects = self.n, self.nc, len(self.p_body), len(self.rects)
        d.inv_mass, d.inv_inertia = f(self.im), f(self.ii)
        d.c_type, d.c_a, d.c_b, d.c_owner = i(self.type), i(self.a), i(self.b), i(self.owner)
        d.c_lax, d.c_lay, d.c_lbx, d.c_lby = f(self.lax), f(self.lay), f(self.lbx), f(self.lby)
        d.c_rest, d.c_min, d.c_max = f(self.rest), f(self.minA), f(self.maxA)
        d.c_color = i(self.color)
        d.p_body, d.p_lx, d.p_ly, d.p_r = i(self.p_body), f(self.p_lx), f(self.p_ly), f(self.p_r)
        d.rects = f(self.rects.reshape(-1))
        c = K.CreatureDesc()
        c.n_creatures, c.n_arms = self.n_creatures, len(self.arm_tip)
        c.creature_arm_off, c.arm_tip, c.arm_grip = i(self.arm_off), i(self.arm_tip), i(self.arm_grip)
        c.arm_motor_off, c.motors = i(self.motor_off), i(self.motors)
        return d, c


class Builder:
    def __init__(self):
        self.p = K.lib().kx_builder_new()
        self.spawns = []
        self.map_args = None

    def __del__(self):
        if getattr(self, "p", None):
            K.lib().kx_builder_free(self.p)
            self.p = None

    def map(self, width, height, seed=1, cols=4, rows=3, density=0.48, wall=14.0):
        self.width, self.height = width, height
        self.map_args = (width, height, seed, cols, rows)
        return K.check(K.lib().kx_builder_map(self.p, width, height, cols, rows, density, wall, seed & 0xFFFFFFFF))

    def find_spawn(self, clearance, rng_state):
        st = C.c_uint32(rng_state[0])
        xy = np.zeros(2)
        K.lib().kx_builder_find_spawn(self.p, clearance, C.byref(st), K.ptr(xy, C.c_double))
        rng_state[0] = st.value
        return xy

    def crn_spawns(self, pop, rng_state, pair_shared=True):
        """crnSpawns (src/sim/train.js:103-122) on this builder's map from the
        trainer's stream state rng_state[0] (advanced in place): [pop, 3]
        x, y, heading."""
        st = C.c_uint32(rng_state[0])
        out = np.zeros(3 * max(pop, 1))
        K.check(K.lib().kx_builder_crn_spawns(self.p, pop, 1 if pair_shared else 0, C.byref(st),
                                              K.ptr(out, C.c_double)))
        rng_state[0] = st.value
        return out[: 3 * pop].reshape(pop, 3)

    def spawn(self, x, y, angle, arms=6, segs=7, ring=8, curl=0.28):
        self.spawns.append((x, y, angle, arms, segs, ring, curl))
        return K.check(K.lib().kx_builder_spawn(self.p, x, y, angle, arms, segs, ring, curl))

    def world(self, solver=None):
        w = World(self.p, self.width, self.height, solver)
        w.spawns = list(self.spawns)
        w.map_seed = self.map_args
        return w


def _rng_next(state):
    r = mulberry32(state[0])
    v = r()
    state[0] = r.state[0]
    return v


def world_cpu64(n=64, seed=1, solver=None):
    """Config 0: createSim-style spawns (core.js:37-58) on makeMap(1200,800)."""
    b = Builder()
    b.map(1200, 800, seed)
    st = [(seed ^ 0x9E3779B9) & 0xFFFFFFFF]
    for _ in range(n):
        xy = b.find_spawn(ARM_REACH + 20, st)
        ang = _rng_next(st) * math.pi * 2
        b.spawn(xy[0], xy[1], ang)
    return b.world(solver)


def world_uniform(n=4096, seed=1, jitter=20.0, solver=None):
    """Config 1: one createSim base pose, +-jitter px per creature."""
    b = Builder()
    b.map(1200, 800, seed)
    st = [(seed ^ 0x9E3779B9) & 0xFFFFFFFF]
    xy = b.find_spawn(ARM_REACH + 20, st)
    ang = _rng_next(st) * math.pi * 2
    r = mulberry32(seed ^ 0x5BF03635)
    for _ in range(n):
        dx = (r() * 2 - 1) * jitter
        dy = (r() * 2 - 1) * jitter
        b.spawn(xy[0] + dx, xy[1] + dy, ang)
    return b.world(solver)


def morphology(seed, i):
    """Config 2 morphology draw: arms 4-8, segments 5-9, mantle ring 6-10."""
    r = mulberry32((seed * 2654435761 + i * 0x9E3779B9) & 0xFFFFFFFF)
    return 4 + int(r() * 5), 5 + int(r() * 5), 6 + int(r() * 5)


def world_uneven(n=8192, seed=1, solver=None, sort_by_size=True):
    """Config 2: seeded uneven morphologies on makeMap(1200,800)."""
    b = Builder()
    b.map(1200, 800, seed)
    st = [(seed ^ 0x9E3779B9) & 0xFFFFFFFF]
    morph = [morphology(seed, i) for i in range(n)]
    order = sorted(range(n), key=lambda i: (morph[i][0] * morph[i][1], morph[i])) if sort_by_size else range(n)
    spawns = []
    for _ in range(min(n, 256)):
        xy = b.find_spawn(ARM_REACH + 20, st)
        spawns.append((xy[0], xy[1], _rng_next(st) * math.pi * 2))
    for j, i in enumerate(order):
        arms, segs, ring = morph[i]
        x, y, a = spawns[j % len(spawns)]
        b.spawn(x, y, a, arms, segs, ring)
    w = b.world(solver)
    w.morph = [morph[i] for i in order]
    return w


def world_from_spawns(spawns, map_args, solver=None):
    """A world of the given spawn tuples (x, y, angle, arms, segs, ring, curl)
    on the map built from map_args (width, height, seed, cols, rows)."""
    b = Builder()
    b.map(*map_args)
    for s in spawns:
        b.spawn(*s)
    return b.world(solver)


def trainer_stream(seed):
    """The trainer's stream state (src/sim/train.js:65-72): seed*2654435761."""
    return [(seed * 2654435761) & 0xFFFFFFFF]


def world_crn(pop, seed=1, map_seed=1, width=1200, height=800, arms=6, segs=7, ring=8, solver=None):
    """An ES population on crnSpawns (src/sim/train.js:103-122): mirror pairs
    share spawn and heading; returns the world and the spawn table."""
    b = Builder()
    b.map(width, height, map_seed)
    sp = b.crn_spawns(pop, trainer_stream(seed), True)
    for x, y, a in sp:
        b.spawn(x, y, a, arms, segs, ring)
    w = b.world(solver)
    w.crn = sp
    return w, sp


def world_stress(n=65536, seed=1, solver=None):
    """Config 4: makeMap(4096,4096,{32x32 corridors}); creatures start in the
    corridor cells (cell centre +-20 px) so arms reach the walls."""
    b = Builder()
    b.map(4096, 4096, seed, 32, 32)
    r = mulberry32(seed ^ 0x5BF03635)
    cs = 4096 / 32
    for i in range(n):
        c = i % 1024
        x = (c % 32 + 0.5) * cs + (r() * 2 - 1) * 20
        y = (c // 32 + 0.5) * cs + (r() * 2 - 1) * 20
        b.spawn(x, y, r() * math.pi * 2)
    return b.world(solver)


def default_law(seed=2024, tone=0.5, grip_min=40.0, grip_max=900.0):
    """The seeded control law (per-motor targets, adhesion reflex) with the
    reference's constants (kraken.js:22-47, core.js:30-31, 183-220)."""
    L = K.Law()
    L.seed = seed & 0xFFFFFFFF
    L.tone = tone
    L.shoulder_limit = SHOULDER_LIMIT
    L.joint_limit = JOINT_LIMIT
    L.stiff_min, L.stiff_max = 2000.0, 400000.0
    L.grip_min, L.grip_max = grip_min, grip_max
    L.slip_max, L.cooldown = 150.0, 0.05
    L.attach_min, L.release_below = 0.08, 0.04
    return L
