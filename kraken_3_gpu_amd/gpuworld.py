"""The gpuworld surface over libkxpbd (Python host).

Mirrors src/gpu/gpuworld.js function for function — create, upload,
uploadControls, packControls, step, download, mirrorInto, groupCollPoints,
destroy, GPU_STRIDES — with the same argument meaning and error behaviour
(thrown errors, NaN counted not raised), so a caller written against the
reference's Backend interface (src/sim/gputrain.js:31-47) runs unchanged:
`backend = gpu_backend(g)` gives upload/uploadControls/step/download.

The device layout is SoA; the 48-byte body / 80-byte constraint AoS records
exist only in the pure helpers pack_controls / mirror_into, which keep the
reference's offsets (body 12 floats, constraint 20 words: lbx o+6, lby o+7,
compliance o+8, target o+10, breakForce o+16, broken o+17 as u32,
lambdaX/Y o+13/14, lambda o+15).
"""
import ctypes as C

import numpy as np

from . import _lib as K

GPU_STRIDES = {"body": 12, "constraint": 20, "collPt": 4, "rect": 4, "force": 4}
WORKGROUP = 64


def v8_hypot(x, y):
    """Math.hypot for two arguments, vectorised, operation for operation as
    V8 evaluates it (normalised Kahan sum), so grip forces mirror bit-exactly."""
    x = np.abs(np.asarray(x, dtype=np.float64))
    y = np.abs(np.asarray(y, dtype=np.float64))
    m = np.maximum(x, y)
    with np.errstate(invalid="ignore", divide="ignore"):
        n0, n1 = x / m, y / m
        s, comp = 0.0, 0.0
        out_s = np.zeros_like(m)
        out_c = np.zeros_like(m)
        for n in (n0, n1):
            summand = n * n - out_c
            prelim = out_s + summand
            out_c = (prelim - out_s) - summand
            out_s = prelim
        r = np.sqrt(out_s) * m
    r = np.where(m == 0, 0.0, r)
    r = np.where(np.isnan(x) | np.isnan(y), np.nan, r)
    r = np.where(np.isinf(x) | np.isinf(y), np.inf, r)
    return r


class Handle:
    """gpuworld handle `g`: owns the device world; keeps a reference to the
    caller's World (`g.cpu`) like gpuworld.create does."""

    def __init__(self, cpu, path, lanes_per_creature, waves_per_block):
        self.cpu = cpu
        self.p = C.c_void_p()
        keep = []
        d, c = cpu.desc(keep)
        prm = K.Params()
        s = cpu.solver
        prm.substeps, prm.iterations, prm.h = s["substeps"], s["iterations"], s["h"]
        prm.lin_drag, prm.ang_drag = s["lin_drag"], s["ang_drag"]
        prm.max_speed, prm.max_ang_speed = s["max_speed"], s["max_ang_speed"]
        prm.path, prm.lanes_per_creature, prm.waves_per_block = path, lanes_per_creature, waves_per_block
        has_cr = cpu.n_creatures > 0 and len(cpu.arm_tip) > 0
        K.check(K.lib().kx_create(C.byref(d), C.byref(c) if has_cr else None, C.byref(prm), C.byref(self.p)))
        nb, nc, npt, ncr = (C.c_int32() for _ in range(4))
        K.check(K.lib().kx_counts(self.p, C.byref(nb), C.byref(nc), C.byref(npt), C.byref(ncr)))
        self.nB, self.nC, self.nP, self.nCr = nb.value, nc.value, npt.value, ncr.value
        self.nR = len(cpu.rects)
        order = np.zeros(self.nC, dtype=np.int32)
        color = np.zeros(self.nC, dtype=np.int32)
        K.check(K.lib().kx_constraint_order(self.p, K.ptr(order, C.c_int32), K.ptr(color, C.c_int32)))
        # ensureColored semantics on the caller's world: colour + stable sort
        cpu.color = cpu.color.copy()
        cpu.color[order] = color
        cpu.reorder_constraints(order)
        off = (C.c_uint32 * 16)()
        cnt = (C.c_uint32 * 16)()
        K.check(K.lib().kx_color_info(self.p, off, cnt))
        cpu.colorOff, cpu.colorCnt = np.array(off, dtype=np.uint32), np.array(cnt, dtype=np.uint32)
        self.colorWG = max(1, -(-int(cpu.colorCnt.max()) // WORKGROUP))
        self.dispatches = 0
        self.path = path
        self._f32 = {}

    def _buf(self, name, n, dt=np.float32):
        a = self._f32.get(name)
        if a is None or len(a) != n:
            a = np.zeros(n, dtype=dt)
            self._f32[name] = a
        return a

    def close(self):
        if self.p:
            K.lib().kx_destroy(self.p)
            self.p = None

    def __del__(self):
        self.close()


def is_available():
    try:
        K.lib()
    except ImportError:
        return False
    return True


def create(cpu, path=K.KX_PATH_PERSISTENT, lanes_per_creature=0, waves_per_block=0):
    """gpuworld.create (gpuworld.js:34-171): colours the world, builds the
    device buffers; raises on implausible params or bad indices."""
    return Handle(cpu, path, lanes_per_creature, waves_per_block)


def _cons_state(g, cpu, with_dynamics):
    f = lambda name, v: _fill(g, name, v)
    cs = K.ConsState()
    cs.lbx, cs.lby = f("lbx", cpu.lbx), f("lby", cpu.lby)
    cs.compliance, cs.target = f("comp", cpu.compliance), f("tgt", cpu.target)
    bf = np.where(np.isfinite(cpu.breakForce), np.minimum(cpu.breakForce, 1e30), 1e30)
    cs.break_force = f("bf", bf)
    br = g._buf("broken", g.nC, np.uint8)
    br[:] = cpu.broken != 0
    cs.broken = K.ptr(br, C.c_uint8)
    if with_dynamics:
        cs.lambda_, cs.lambda_x, cs.lambda_y = f("lam", cpu.lambda_), f("lamx", cpu.lambdaX), f("lamy", cpu.lambdaY)
    return cs


def _fill(g, name, v):
    a = g._buf(name, len(v))
    a[:] = v
    return K.ptr(a, C.c_float)


def upload(g):
    """Full CPU state -> device (gpuworld.js:250-298)."""
    cpu = g.cpu
    bs = K.BodyState()
    f = lambda name, v: _fill(g, "b_" + name, v)
    bs.px, bs.py, bs.vx, bs.vy = f("px", cpu.px), f("py", cpu.py), f("vx", cpu.vx), f("vy", cpu.vy)
    bs.ang, bs.av = f("ang", cpu.ang), f("av", cpu.av)
    bs.fx, bs.fy, bs.tq = f("fx", cpu.fx), f("fy", cpu.fy), f("tq", cpu.tq)
    cs = _cons_state(g, cpu, True)
    K.check(K.lib().kx_upload(g.p, C.byref(bs), C.byref(cs)))
    cd = np.ascontiguousarray(cpu.cd, dtype=np.float64)
    K.check(K.lib().kx_set_cooldown(g.p, K.ptr(cd, C.c_double)))


def pack_controls(cpu, cons_data):
    """packControls (gpuworld.js:331-341), pure: writes lbx/lby (o+6/7),
    compliance (o+8), target (o+10), breakForce (o+16, inf -> 1e30) and
    broken (o+17, through the u32 view) into a float32 AoS block."""
    n = cpu.nc
    rec = cons_data[: n * 20].reshape(n, 20)
    rec[:, 6] = cpu.lbx
    rec[:, 7] = cpu.lby
    rec[:, 8] = cpu.compliance
    rec[:, 10] = cpu.target
    rec[:, 16] = np.where(np.isfinite(cpu.breakForce), np.minimum(cpu.breakForce, 1e30), 1e30)
    cons_data.view(np.uint32)[: n * 20].reshape(n, 20)[:, 17] = (cpu.broken != 0).astype(np.uint32)
    return cons_data


def upload_controls(g):
    """uploadControls (gpuworld.js:312-315): the control fields only."""
    cs = _cons_state(g, g.cpu, False)
    K.check(K.lib().kx_pack_controls(g.p, C.byref(cs)))


def step(g, n=1):
    """One (or n) simulation steps: S substeps + grip break each, queued."""
    K.check(K.lib().kx_step(g.p, n))
    g.dispatches += n * (1 + g.cpu.S * (3 + g.cpu.I * (int((g.cpu.colorCnt > 0).sum()) + 1)))


def substep(g, n=1):
    """One (or n) substeps without the grip break (gpuworld.js:344-372)."""
    K.check(K.lib().kx_substep(g.p, n))
    g.dispatches += n * (3 + g.cpu.I * (int((g.cpu.colorCnt > 0).sum()) + 1))


def download(g):
    """Device -> CPU world, returns {maxSpeed, maxAngSpeed, nan}
    (gpuworld.js:385-396 + mirrorInto 417-456)."""
    cpu = g.cpu
    nb, nc = g.nB, g.nC
    b = {k: g._buf("d_" + k, nb) for k in ("px", "py", "vx", "vy", "ang", "av")}
    bs = K.BodyState()
    for k, a in b.items():
        setattr(bs, k, K.ptr(a, C.c_float))
    c = {k: g._buf("e_" + k, nc) for k in ("lam", "lamx", "lamy")}
    br = g._buf("e_broken", nc, np.uint8)
    cs = K.ConsState()
    cs.lambda_, cs.lambda_x, cs.lambda_y = (K.ptr(c[k], C.c_float) for k in ("lam", "lamx", "lamy"))
    cs.broken = K.ptr(br, C.c_uint8)
    st = K.Stats()
    K.check(K.lib().kx_download(g.p, C.byref(bs), C.byref(cs), C.byref(st)))
    cpu.px[:], cpu.py[:], cpu.vx[:], cpu.vy[:] = b["px"], b["py"], b["vx"], b["vy"]
    cpu.ang[:], cpu.av[:] = b["ang"], b["av"]
    cpu.lambda_[:] = c["lam"]
    cpu.broken[:] = br
    grip = cpu.type == K.KX_GRIP
    cpu.lambdaX[grip] = c["lamx"][grip]
    cpu.lambdaY[grip] = c["lamy"][grip]
    cpu.force[grip] = v8_hypot(cpu.lambdaX[grip], cpu.lambdaY[grip]) * (1 / cpu.h)
    stats = dict(maxSpeed=st.max_speed, maxAngSpeed=st.max_ang_speed, nan=int(st.nan))
    cpu.stats.update(stats)
    return stats


def mirror_into(cpu, body_data, cons_data):
    """mirrorInto (gpuworld.js:417-456), pure: AoS float32 blocks -> world."""
    nb, nc = cpu.n, cpu.nc
    bd = body_data[: nb * 12].reshape(nb, 12).astype(np.float64)
    cpu.px[:], cpu.py[:] = bd[:, 0], bd[:, 1]
    cpu.vx[:], cpu.vy[:] = bd[:, 4], bd[:, 5]
    cpu.ang[:], cpu.av[:] = bd[:, 6], bd[:, 8]
    s2 = bd[:, 4] * bd[:, 4] + bd[:, 5] * bd[:, 5]
    with np.errstate(invalid="ignore"):
        max_s2 = float(np.max(np.where(np.isnan(s2), 0.0, s2), initial=0.0))
        max_a = float(np.max(np.where(np.isnan(bd[:, 8]), 0.0, np.abs(bd[:, 8])), initial=0.0))
    fin = np.isfinite(bd[:, [0, 1, 4, 5, 6]]).all(axis=1)
    nan = int((~fin).sum())
    cdat = cons_data[: nc * 20].reshape(nc, 20)
    cu = cons_data.view(np.uint32)[: nc * 20].reshape(nc, 20)
    cpu.lambda_[:] = cdat[:, 15]
    cpu.broken[:] = (cu[:, 17] != 0)
    grip = cpu.type == K.KX_GRIP
    cpu.lambdaX[grip] = cdat[grip, 13]
    cpu.lambdaY[grip] = cdat[grip, 14]
    cpu.force[grip] = v8_hypot(cpu.lambdaX[grip], cpu.lambdaY[grip]) * (1 / cpu.h)
    stats = dict(maxSpeed=float(np.sqrt(max_s2)), maxAngSpeed=max_a, nan=nan)
    cpu.stats.update(stats)
    return stats


def group_coll_points(p_body, n):
    """groupCollPoints (gpuworld.js:189-207): (order, range) with range =
    offsets then counts; raises on a bad body index."""
    p_body = np.ascontiguousarray(p_body, dtype=np.int32)
    order = np.zeros(len(p_body), dtype=np.int32)
    rng = np.zeros(2 * n, dtype=np.int32)
    K.check(K.lib().kx_group_points(len(p_body), K.ptr(p_body, C.c_int32), n, K.ptr(order, C.c_int32),
                                    K.ptr(rng, C.c_int32)))
    return order, rng


def destroy(g):
    """Idempotent release of the device world (gpuworld.js:466-472)."""
    if g is not None:
        g.close()


def rollout(g, law, step0, nsteps, events=False):
    """nsteps of device-side control law + reflex + physics; returns the
    per-step event words ([nsteps, n_creatures] u32) when asked."""
    ev = np.zeros((nsteps, g.nCr), dtype=np.uint32) if events else None
    K.check(K.lib().kx_rollout(g.p, C.byref(law), step0, nsteps, K.ptr(ev, C.c_uint32) if events else None))
    return ev


def upload_tape(g, tape):
    """Control tape for KX_LAW_TAPE rollouts: float32 [rows, n_creatures,
    KX_TAPE_WORDS] (tone, then shoulder, curl, adhesion per arm)."""
    t = np.ascontiguousarray(tape, dtype=np.float32)
    if t.ndim != 3 or t.shape[1] != g.nCr or t.shape[2] != K.KX_TAPE_WORDS:
        raise ValueError(f"tape shape {t.shape}, want (rows, {g.nCr}, {K.KX_TAPE_WORDS})")
    K.check(K.lib().kx_upload_tape(g.p, K.ptr(t, C.c_float), t.shape[0]))


def rollout_graph(g, law, step0, nsteps, reps):
    K.check(K.lib().kx_rollout_graph(g.p, C.byref(law), step0, nsteps, reps))


def sync(g):
    K.check(K.lib().kx_sync(g.p))


def last_timing(g):
    ms = C.c_double()
    n = C.c_int32()
    K.check(K.lib().kx_last_timing(g.p, C.byref(ms), C.byref(n)))
    return ms.value, n.value


def set_path(g, path, lanes_per_creature=0, waves_per_block=0):
    K.check(K.lib().kx_set_path(g.p, path, lanes_per_creature, waves_per_block))
    g.path = path


def schedule_rounds(g):
    """Rounds per iteration of every morphology template (persistent path)."""
    buf = np.zeros(4096, dtype=np.int32)
    n = K.check(K.lib().kx_schedule_info(g.p, K.ptr(buf, C.c_int32), len(buf)))
    return buf[:n].tolist()


PERSIST_INFO = ("lds_bytes", "workgroups", "resident_per_cu", "needed_per_cu", "map_words", "bitmap_words",
                "template_words", "creature_words", "mixed_lanes", "collision_fold", "register_rounds", "fixed_rounds",
                "fused_creatures")


def set_register_slots(g, enable):
    """The persistent kernels specialised to the default creature's schedule
    (kx_set_register_slots): True or 1 (the default) register-held round
    slots where the launch allows them, else slots read from LDS; 2 the
    LDS-read form always; False or 0 neither. Results are identical."""
    mode = 1 if enable is True else 0 if enable is False else int(enable)
    K.check(K.lib().kx_set_register_slots(g.p, mode))


def persist_info(g):
    """Launch shape of the persistent path for the handle's (lanes, waves)."""
    buf = np.zeros(len(PERSIST_INFO), dtype=np.int32)
    K.check(K.lib().kx_persist_info(g.p, K.ptr(buf, C.c_int32), len(buf)))
    return dict(zip(PERSIST_INFO, buf.tolist()))


STATE_EXTRA = ("x", "y", "angle", "vx", "vy", "av", "speed", "mean_extension", "mean_curl", "anchored", "ax", "ay")


def read_state(g, stride=0):
    """Per-creature state readout (kx_read_state): readProprio's 6*A + 6
    inputs (src/sim/core.js:129-175) followed by the mantle state
    STATE_EXTRA, float32 [nCr, stride]. Advances the previous mantle velocity
    the acceleration inputs use, as readProprio does once per step."""
    off = np.asarray(g.cpu.arm_off)
    arms = int(np.diff(off).max()) if len(off) > 1 else 0
    need = 6 * arms + 6 + K.KX_STATE_EXTRA
    stride = max(stride, need)
    out = np.zeros((g.nCr, stride), np.float32)
    K.check(K.lib().kx_read_state(g.p, K.ptr(out, C.c_float), stride))
    return out


def set_collision_fold(g, fold):
    """Persistent path: collision fold (K.KX_FOLD_AUTO, _PARALLEL or _OWNER);
    every fold gives the same result."""
    K.check(K.lib().kx_set_collision_fold(g.p, fold))


TEMPLATE_INFO = ("bodies", "constraints", "points", "creatures", "rounds", "lanes")


def template_info(g):
    """Per morphology template: size, population, rounds and lane count."""
    n = K.check(K.lib().kx_template_info(g.p, -1, None))
    out = []
    buf = np.zeros(6, dtype=np.int32)
    for t in range(n):
        K.check(K.lib().kx_template_info(g.p, t, K.ptr(buf, C.c_int32)))
        out.append(dict(zip(TEMPLATE_INFO, buf.tolist())))
    return out


def creature_templates(g):
    """Template index of every creature."""
    out = np.zeros(g.nCr, dtype=np.int32)
    K.check(K.lib().kx_creature_templates(g.p, K.ptr(out, C.c_int32)))
    return out


def set_template_lanes(g, lanes):
    """Lanes per creature of every template (0 = the handle's default)."""
    a = np.ascontiguousarray(lanes, dtype=np.int32)
    K.check(K.lib().kx_set_template_lanes(g.p, len(a), K.ptr(a, C.c_int32)))


def set_plain_sdf(g, use_grid):
    K.check(K.lib().kx_set_plain_sdf(g.p, 1 if use_grid else 0))


def respawn(g, creatures, poses):
    """respawnCreature (src/sim/train.js:188-229) in place on the device for
    the listed creatures (poses [n, 3]: x, y, heading); the caller's world
    is not touched until the next download."""
    ids = np.ascontiguousarray(creatures, dtype=np.int32)
    pz = np.ascontiguousarray(poses, dtype=np.float64).reshape(-1, 3)
    if len(pz) != len(ids):
        raise ValueError(f"{len(ids)} creatures but {len(pz)} poses")
    K.check(K.lib().kx_respawn(g.p, len(ids), K.ptr(ids, C.c_int32), K.ptr(pz, C.c_double)))


def download_cooldown(g):
    cd = np.zeros((g.nCr, 8))
    K.check(K.lib().kx_get_cooldown(g.p, K.ptr(cd, C.c_double)))
    g.cpu.cd[:] = cd
    return cd


def download_controls(g):
    """lbx, lby, compliance, target, breakForce of every constraint (handle order)."""
    nc = g.nC
    out = {k: np.zeros(nc, dtype=np.float32) for k in ("lbx", "lby", "compliance", "target", "breakForce")}
    cs = K.ConsState()
    cs.lbx, cs.lby = K.ptr(out["lbx"], C.c_float), K.ptr(out["lby"], C.c_float)
    cs.compliance, cs.target = K.ptr(out["compliance"], C.c_float), K.ptr(out["target"], C.c_float)
    cs.break_force = K.ptr(out["breakForce"], C.c_float)
    K.check(K.lib().kx_download(g.p, None, C.byref(cs), None))
    return out


def gpu_backend(g):
    """Backend object of src/sim/gputrain.js:39-47."""
    class Backend:
        kind = "gpu"

        @staticmethod
        def upload():
            upload(g)

        @staticmethod
        def uploadControls():
            upload_controls(g)

        @staticmethod
        def step():
            step(g)

        @staticmethod
        def download():
            return download(g)
    return Backend()
