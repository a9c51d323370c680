"""The reference's own GPU acceptance gates (gpu.html G1-G3, G7) restated on
the HIP paths, plus the respawn, non-finite-state and SDF-fold checks.

* G1 passive rest for every uneven shape (gpu.html:223-268),
* G2 spawn violation after upload + one step < 1e-3 px (gpu.html:271-300),
* G3 grip-and-pull within +-30 % of the CPU oracle (gpu.html:303-339),
* G7 lock-step (gpu.html:478-546): 32 creatures on crnSpawns for 150 steps
  under one control law, device vs the f64 oracle AND vs the f32 emulator of
  the reference's WGSL solver (tools/f32solver.mjs): mantle max delta < 5 px,
  grip flags >= 99.5 % identical, broken-grip steps within 10 %, mean path
  length within 5 %,
* respawnCreature in place (src/sim/train.js:188-229) against the oracle, and
  a respawned population against a freshly built one,
* NaN / Inf in the state against the reference's guards (xpbd.js:280-281,
  350, 361, 407),
* the SDF fold (map.js:91-118, best^2 quirk included) through every device
  collision path against the oracle's fold.
"""
import ctypes as C
import math
import sys

import numpy as np
import pytest

from helpers import creature_of_bodies, creature_ranges, load_state, oracle_twin
from kraken_3_gpu_amd import _lib as K
from kraken_3_gpu_amd import gpuworld as G
from kraken_3_gpu_amd import world as W
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

PATHS = {"plain": K.KX_PATH_PLAIN, "persistent": K.KX_PATH_PERSISTENT}
ARM_REACH = 24 + 7 * 13.5
SHAPES = [(a, s, r) for a in range(4, 9) for s in range(5, 10) for r in range(6, 11)]
REF_SOLVER = dict(substeps=2, iterations=4, h=1 / 120)     # xpbd.js:46-54 defaults, gpu.html's buildWorld


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _f32(a):
    return np.asarray(a, np.float32).astype(np.float64)


def _build_world(n, seed, solver=REF_SOLVER):
    """gpu.html buildWorld (gpu.html:200-211): makeMap(1440, 810, {seed}),
    findSpawn(ARM_REACH + 20) from mulberry32(seed ^ 0x5bf03635), heading
    rnd() * 2 pi."""
    b = W.Builder()
    b.map(1440, 810, seed)
    st = [(seed ^ 0x5BF03635) & 0xFFFFFFFF]
    for _ in range(n):
        xy = b.find_spawn(ARM_REACH + 20, st)
        r = W.mulberry32(st[0])
        ang = r() * math.pi * 2
        st[0] = r.state[0]
        b.spawn(xy[0], xy[1], ang)
    return b.world(solver)


def _pin_violation(w):
    worst = 0.0
    for k in np.nonzero((w.type == K.KX_PIN) & (w.broken == 0))[0]:
        A, B = int(w.a[k]), int(w.b[k])
        ca, sa = math.cos(w.ang[A]), math.sin(w.ang[A])
        gax = w.px[A] + w.lax[k] * ca - w.lay[k] * sa
        gay = w.py[A] + w.lax[k] * sa + w.lay[k] * ca
        cb, sb = math.cos(w.ang[B]), math.sin(w.ang[B])
        gbx = w.px[B] + w.lbx[k] * cb - w.lby[k] * sb
        gby = w.py[B] + w.lbx[k] * sb + w.lby[k] * cb
        worst = max(worst, math.hypot(gax - gbx, gay - gby))
    return worst


# ── G1 for every shape ──

@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_g1_passive_rest_every_uneven_shape(path):
    """All 125 shapes of the uneven workload (4-8 arms x 5-9 segments, ring
    6-10), no drive and no contact: 240 steps leave every binary32 pose
    bit-identical and every velocity 0 (gpu.html G1: 'Soll exakt 0')."""
    b = W.Builder()
    b.map(1200, 800, 1, 1, 1)
    for i, (A, S, R) in enumerate(SHAPES):
        b.spawn(600.0, 400.0, 0.05 * i, A, S, R)
    w = b.world()
    g = G.create(w, path=PATHS[path])
    G.upload(g)
    p0 = [np.float32(getattr(w, k)).copy() for k in ("px", "py", "ang")]
    say(f"[case] G1 passive {len(SHAPES)} shapes path={path}")
    G.step(g, 240)
    st = G.download(g)
    for k, a in zip(("px", "py", "ang"), p0):
        assert np.array_equal(np.float32(getattr(w, k)), a), k
    assert not np.any(w.vx) and not np.any(w.vy) and not np.any(w.av)
    assert st["maxSpeed"] == 0.0 and st["maxAngSpeed"] == 0.0 and st["nan"] == 0
    G.destroy(g)


# ── G2 ──

@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_g2_spawn_violation_after_one_step(path):
    w = _build_world(8, 23)
    before = _pin_violation(w)
    g = G.create(w, path=PATHS[path])
    G.upload(g)
    say(f"[case] G2 spawn violation path={path}")
    G.step(g)
    G.download(g)
    after = _pin_violation(w)
    say(f"    G2 pin violation: spawn {before:.2e} px, after upload + 1 step {after:.2e} px")
    # the spawn pose in float64 against the binary32 anchor offsets the
    # device holds: a few 1e-7 px; the gate is the reference's 1e-3
    assert before < 1e-5 and after < 1e-3
    G.destroy(g)


# ── G3 ──

def _g3_setup(w, o):
    """gpu.html:307-325: arm 0 holds at its tip with breakForce 900
    (adhesion 1), its motors stiff (tone 1) curling to 0.98 of the joint
    limit, the other arms slack (tone 0.2) at target 0."""
    o.attach_grip(0, 0, 900.0)
    c = o.cons()
    jl = math.pi * 0.42
    arm0 = w.motors[w.motor_off[0]:w.motor_off[1]]
    comp = np.where(w.type == K.KX_ANGLE, 1 / (2000 + 0.2 * 398000), c["compliance"])
    tgt = np.where(w.type == K.KX_ANGLE, 0.0, c["target"])
    comp[arm0] = 1 / (2000 + 1.0 * 398000)
    tgt[arm0[1:]] = jl * 0.98
    tgt[arm0[0]] = 0.0
    o.set_cons_dynamic(compliance=comp, target=tgt)
    c = o.cons()
    w.compliance[:], w.target[:] = _f32(c["compliance"]), _f32(c["target"])
    w.lbx[:], w.lby[:], w.breakForce[:] = _f32(c["lbx"]), _f32(c["lby"]), _f32(c["breakForce"])
    w.broken[:] = c["broken"].astype(np.uint8)
    load_state(o, w)                   # identical binary32 controls on both sides
    gk = int(w.arm_grip[0])
    return (c["lbx"][gk], c["lby"][gk]), gk


@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_g3_grip_and_pull_matches_cpu(path):
    w = _build_world(1, 31)
    o = oracle_twin(w)
    g = G.create(w, path=PATHS[path])
    anchor, gk = _g3_setup(w, o)
    G.upload(g)
    m = 0
    start = math.hypot(w.px[m] - anchor[0], w.py[m] - anchor[1])
    res = {}
    say(f"[case] G3 grip-and-pull path={path}")
    for side in ("cpu", "gpu"):
        mind, broke = start, False
        for _ in range(600):
            if side == "cpu":
                o.step()
                b = o.bodies()
                x, y, broke = b["px"][m], b["py"][m], bool(o.cons()["broken"][gk])
            else:
                G.step(g)
                G.download(g)
                x, y, broke = w.px[m], w.py[m], bool(w.broken[gk])
            if broke:
                break
            mind = min(mind, math.hypot(x - anchor[0], y - anchor[1]))
        res[side] = (start - mind, broke)
    (pc, bc), (pg, bg) = res["cpu"], res["gpu"]
    rel = abs(pg - pc) / max(1.0, pc)
    say(f"    G3 pulled cpu {pc:.2f} px, gpu {pg:.2f} px ({100 * rel:.2f} %), broken {bc}/{bg}")
    assert pc > 10.0
    assert rel < 0.30 and pg > 0.5 * pc and bc == bg
    G.destroy(g)


# ── G7 ──

def _lockstep(o, g, w, seed, steps):
    """Per step: mantle max delta, grip flag mismatches, broken-grip counts,
    mantle path lengths; device vs oracle world o under the same law."""
    ncr = g.nCr
    mantles = np.array([lo for lo, _ in creature_ranges(w)])
    grips_w = np.asarray(w.arm_grip)
    grips_o = np.array([o.grip_pos(c, a) for c in range(ncr) for a in range(6)])
    prev_w = np.c_[w.px[mantles], w.py[mantles]]
    prev_o = prev_w.copy()
    path_w = np.zeros(ncr)
    path_o = np.zeros(ncr)
    max_d = mism = tot = brk_w = brk_o = 0
    first_flip = -1
    for s in range(steps):
        o.apply_law(seed, s)
        o.step()
        G.rollout(g, W.default_law(seed), s, 1)
        G.download(g)
        b = o.bodies()
        pw = np.c_[w.px[mantles], w.py[mantles]]
        po_ = np.c_[b["px"][mantles], b["py"][mantles]]
        max_d = max(max_d, float(np.hypot(*(pw - po_).T).max()))
        path_w += np.hypot(*(pw - prev_w).T)
        path_o += np.hypot(*(po_ - prev_o).T)
        prev_w, prev_o = pw, po_
        fw = w.broken[grips_w].astype(bool)
        fo = o.cons()["broken"][grips_o].astype(bool)
        d = int((fw != fo).sum())
        if d and first_flip < 0:
            first_flip = s
        mism += d
        tot += fw.size
        brk_w += int(fw.sum())
        brk_o += int(fo.sum())
    pl_w, pl_o = float(path_w.mean()), float(path_o.mean())
    return dict(max_d=max_d, flags=1 - mism / tot, brk_o=brk_o, brk_w=brk_w, pl_o=pl_o, pl_w=pl_w,
                pl_dev=100 * (pl_w - pl_o) / max(1e-9, abs(pl_o)), first_flip=first_flip)


def _oracle_pair_spread(mk, seed, steps):
    """The same G7 measures between the f32 emulator of the reference's WGSL
    solver and the f64 oracle: how far the reference's own GPU arithmetic
    drifts from f64 on this workload."""
    w = mk()
    a, b = oracle_twin(w), oracle_twin(w)
    b.f32ify()
    m = np.array([lo for lo, _ in creature_ranges(w)])
    ga = np.array([a.grip_pos(c, k) for c in range(w.n_creatures) for k in range(6)])
    gb = np.array([b.grip_pos(c, k) for c in range(w.n_creatures) for k in range(6)])
    max_d, mism, tot = 0.0, 0, 0
    for s in range(steps):
        for o in (a, b):
            o.apply_law(seed, s)
            o.step()
        A, B = a.bodies(), b.bodies()
        max_d = max(max_d, float(np.hypot(A["px"][m] - B["px"][m], A["py"][m] - B["py"][m]).max()))
        fa, fb = a.cons()["broken"][ga], b.cons()["broken"][gb]
        mism += int((fa != fb).sum())
        tot += fa.size
    return dict(max_d=max_d, flags=1 - mism / tot)


@pytest.mark.parametrize("setting", ["reference", "config"])
@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_g7_lockstep_against_f64_and_f32_emulator(path, setting):
    """createTrainer's population (gpu.html:519-522): 32 creatures on
    crnSpawns, pairs sharing spawn and heading; 150 steps of the seeded law
    on the device and on each oracle.

    reference: the reference's solver default (S=2, I=4, h=1/120 on
    makeMap(1440, 810, {seed: 7})), held to G7's own gates.
    config: the benchmark's S=4, I=8, h=1/240 on makeMap(1200, 800, {seed:
    1}). There the seeded law's violent random targets make single grip
    flips chaotic: the reference's own f32 arithmetic (its WGSL emulator)
    lands 18 px and 98.6 % flags from f64 on this population, outside G7's
    5 px / 99.5 %. The device is held to that envelope instead (no farther
    from either oracle than the emulator is from f64, with a quarter's
    margin), and to G7's break-count and path-length gates unchanged."""
    if setting == "reference":
        solver, (width, height, mseed), seed = REF_SOLVER, (1440, 810, 7), 7
    else:
        solver, (width, height, mseed), seed = None, (1200, 800, 1), 1
    mk = lambda: W.world_crn(32, seed=seed, map_seed=mseed, width=width, height=height, solver=solver)[0]
    rows = {}
    for ref in ("f64", "f32"):
        w = mk()
        o = oracle_twin(w)
        if ref == "f32":
            o.f32ify()
        g = G.create(w, path=PATHS[path])
        G.upload(g)
        say(f"[case] G7 lockstep path={path} setting={setting} vs {ref}")
        r = _lockstep(o, g, w, 2024, 150)
        G.destroy(g)
        rows[ref] = r
        say(f"    G7 vs {ref}: maxD {r['max_d']:.4f} px, flags {100 * r['flags']:.3f} %, broken-steps "
            f"{r['brk_o']} -> {r['brk_w']}, pathLen {r['pl_o']:.2f} -> {r['pl_w']:.2f} ({r['pl_dev']:+.2f} %), "
            f"first flip {r['first_flip']}")
    max_d, flags = 5.0, 0.995
    if setting == "config":
        env = _oracle_pair_spread(mk, 2024, 150)
        say(f"    f32 emulator vs f64 on the same population: maxD {env['max_d']:.4f} px, "
            f"flags {100 * env['flags']:.3f} %")
        max_d = max(max_d, 1.25 * env["max_d"])
        flags = min(flags, env["flags"] - 0.0025)
    for ref, r in rows.items():
        assert r["max_d"] < max_d, ref
        assert r["flags"] >= flags, ref
        assert abs(r["brk_w"] - r["brk_o"]) <= max(3, 0.10 * r["brk_o"]), ref
        assert abs(r["pl_dev"]) < 5.0, ref
    assert rows["f64"]["pl_o"] > 1.0       # the law moves the creatures


# ── respawn ──

@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_respawn_in_place_matches_oracle(path):
    """A population 30 steps into the law; four creatures respawned on the
    device at new crnSpawns draws: their state equals the oracle's
    respawnCreature of the same state rounded to binary32 (poses within one
    binary32 ulp, the rest exact), the other creatures are untouched; then
    one more step within 1e-3 px of the oracle."""
    w, _ = W.world_crn(16, seed=3)
    g = G.create(w, path=PATHS[path])
    G.upload(g)
    law = W.default_law(77)
    G.rollout(g, law, 0, 30)
    G.download(g)
    ctl = G.download_controls(g)
    w.lbx[:], w.lby[:] = ctl["lbx"], ctl["lby"]
    w.compliance[:], w.target[:], w.breakForce[:] = ctl["compliance"], ctl["target"], ctl["breakForce"]
    G.download_cooldown(g)
    before = {k: getattr(w, k).copy() for k in ("px", "py", "ang", "vx", "vy", "av")}
    b2 = W.Builder()
    b2.map(1200, 800, 1)
    poses = b2.crn_spawns(4, W.trainer_stream(99), False)
    ids = [0, 3, 8, 9]
    o = oracle_twin(w)
    load_state(o, w)
    for c, (x, y, a) in zip(ids, poses):
        o.respawn(c, x, y, a)
    say(f"[case] respawn path={path}")
    G.respawn(g, ids, poses)
    G.download(g)
    ob = o.bodies()
    crb = creature_of_bodies(w)
    moved = np.isin(crb, ids)
    for k in ("px", "py", "ang"):
        want = np.float32(ob[k][moved])
        got = np.float32(getattr(w, k)[moved])
        ulps = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
        assert ulps.max() <= 1, k
        assert np.array_equal(getattr(w, k)[~moved], before[k][~moved]), k
    for k in ("vx", "vy", "av"):
        assert not np.any(getattr(w, k)[moved])
        assert np.array_equal(getattr(w, k)[~moved], before[k][~moved]), k
    c = o.cons()
    ctl = G.download_controls(g)
    assert np.array_equal(w.broken, c["broken"].astype(np.uint8))
    assert np.array_equal(ctl["target"], np.float32(c["target"]))
    grips = (w.type == K.KX_GRIP) & np.isin(crb[w.a], ids)
    assert np.all(w.broken[grips] == 1) and not np.any(w.lambdaX[grips]) and not np.any(w.lambdaY[grips])
    # one more step from the respawned state
    w.lbx[:], w.lby[:] = ctl["lbx"], ctl["lby"]
    w.compliance[:], w.target[:], w.breakForce[:] = ctl["compliance"], ctl["target"], ctl["breakForce"]
    G.download_cooldown(g)
    load_state(o, w)
    o.set_step_counter(30)
    o.apply_law(77, 30)
    o.step()
    G.rollout(g, law, 30, 1)
    G.download(g)
    ob = o.bodies()
    err = float(np.maximum(np.abs(w.px - ob["px"]), np.abs(w.py - ob["py"])).max())
    say(f"    respawned 4 of 16; next step max|dpos| {err:.3e} px")
    assert err <= 1e-3
    G.destroy(g)


def test_generation_reset_without_rebuilding_equals_a_fresh_handle():
    """beginGenerationGpu (gputrain.js:130-133) without the host respawn and
    full upload: every creature of a running handle respawned at its
    crnSpawns slot. Its poses equal a freshly built world's binary32 poses
    (to one ulp: device and host sin/cos may differ in the last float64 bit);
    its readout's acceleration restarts from rest; and from the same poses
    (cooldowns cleared, as a fresh world starts them) the next 40 law steps
    equal a freshly built handle's bit for bit, so no other state of the
    previous generation survives the respawn."""
    law = W.default_law(5)
    w1, sp = W.world_crn(64, seed=11)
    g1 = G.create(w1)
    G.upload(g1)
    G.rollout(g1, law, 0, 25)
    G.read_state(g1)
    G.rollout(g1, law, 25, 1)
    say("[case] generation reset: respawn all 64")
    G.respawn(g1, np.arange(64), sp)
    st = G.read_state(g1)
    assert not np.any(st[:, 36:38]), "acceleration inputs after a respawn"
    K.check(K.lib().kx_set_cooldown(g1.p, K.ptr(np.zeros(64 * 8), C.c_double)))
    G.download(g1)
    w2, _ = W.world_crn(64, seed=11)
    exact = 0
    for k in ("px", "py", "ang"):
        a, b = np.float32(getattr(w1, k)), np.float32(getattr(w2, k))
        ulps = np.abs(a.view(np.int32).astype(np.int64) - b.view(np.int32).astype(np.int64))
        assert ulps.max() <= 1, k
        exact += int((ulps == 0).sum())
        getattr(w2, k)[:] = getattr(w1, k)
    g2 = G.create(w2)
    G.upload(g2)
    e1 = G.rollout(g1, law, 0, 40, events=True)
    e2 = G.rollout(g2, law, 0, 40, events=True)
    G.download(g1)
    G.download(g2)
    for k in ("px", "py", "ang", "vx", "vy", "av"):
        assert np.array_equal(getattr(w1, k), getattr(w2, k)), k
    assert np.array_equal(w1.broken, w2.broken) and np.array_equal(e1, e2)
    say(f"    poses {exact}/{3 * w1.n} bit-equal to a fresh build; 40 steps identical, "
        f"{int((e1 != 0).sum())} event words")
    G.destroy(g1)
    G.destroy(g2)


# ── non-finite state ──

@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_non_finite_state_follows_the_reference_guards(path):
    """NaN / Inf planted in the state: the integrate guard zeroes a
    non-finite velocity (xpbd.js:280-281), an infinite spin clamps to the
    limit, a NaN pose stays NaN without spreading (every solve skips a
    non-finite gap, xpbd.js:350, 361, 407), a NaN first body does not shift
    its creature. Over 3 steps, each one from the device's state loaded into
    the oracle: the device's non-finite pattern equals the oracle's and every
    finite body stays within 1e-3 px (the one-step bar)."""
    w = W.world_cpu64(8, seed=2)
    o = oracle_twin(w)
    g = G.create(w, path=PATHS[path])
    G.upload(g)
    law = W.default_law(3)
    G.rollout(g, law, 0, 10)
    G.download(g)
    ctl = G.download_controls(g)
    w.lbx[:], w.lby[:] = ctl["lbx"], ctl["lby"]
    w.compliance[:], w.target[:], w.breakForce[:] = ctl["compliance"], ctl["target"], ctl["breakForce"]
    G.download_cooldown(g)
    first = [lo for lo, _ in creature_ranges(w)]
    w.px[first[1]] = np.nan                  # creature 1: NaN mantle (its origin)
    w.vx[first[2] + 3] = np.nan              # creature 2: NaN velocity
    w.vy[first[3] + 5] = np.inf              # creature 3: infinite velocity
    w.av[first[4] + 2] = -np.inf             # creature 4: infinite spin
    w.ang[first[5] + 7] = np.nan             # creature 5: NaN angle
    w.py[first[6] + 10] = np.inf             # creature 6: infinite position
    G.upload(g)
    load_state(o, w)
    o.set_step_counter(10)
    say(f"[case] non-finite state path={path}")
    for s in range(3):
        if s:
            # the oracle restarts from the device's state (one-step parity)
            ctl = G.download_controls(g)
            w.lbx[:], w.lby[:] = ctl["lbx"], ctl["lby"]
            w.compliance[:], w.target[:], w.breakForce[:] = ctl["compliance"], ctl["target"], ctl["breakForce"]
            G.download_cooldown(g)
            load_state(o, w)
            o.set_step_counter(10 + s)
        o.apply_law(3, 10 + s)
        o.step()
        G.rollout(g, law, 10 + s, 1)
        st = G.download(g)
        b = o.bodies()
        for k in ("px", "py", "ang", "vx", "vy", "av"):
            assert np.array_equal(np.isfinite(getattr(w, k)), np.isfinite(b[k])), (s, k)
        fin = np.isfinite(w.px) & np.isfinite(w.py)
        err = float(np.maximum(np.abs(w.px[fin] - b["px"][fin]), np.abs(w.py[fin] - b["py"][fin])).max())
        nonfin = int((~(np.isfinite(b["px"]) & np.isfinite(b["py"]) & np.isfinite(b["ang"]))).sum())
        say(f"    step {s + 1}: {nonfin} non-finite bodies (device stats nan {st['nan']}), finite max|dpos| {err:.3e}")
        assert err <= 1e-3
        assert st["maxAngSpeed"] <= 18.0 * 1.001
    assert not np.isfinite(w.px[first[1]]) and not np.isfinite(w.py[first[6] + 10])
    G.destroy(g)


# ── SDF fold through the device collision paths ──

def _probe_world(rects, pts, radius):
    """One creature per query point; only the mantle keeps a collision point,
    at its centre with radius `radius`, so one substep x one iteration moves
    the mantle by n * (radius - d) exactly where the fold reports a
    penetration beyond the dead zone."""
    b = W.Builder()
    b.map(1200, 800, 1)
    for x, y in pts:
        b.spawn(float(x), float(y), 0.0, 4, 5, 1)
    w = b.world(dict(substeps=1, iterations=1))
    keep = np.nonzero(np.isin(w.p_body, np.arange(0, w.n, 21)))[0]
    w.p_body, w.p_lx, w.p_ly = w.p_body[keep], np.zeros(len(keep), np.float32), np.zeros(len(keep), np.float32)
    w.p_r = np.full(len(keep), radius, np.float32)
    w.rects = np.asarray(rects, np.float32).reshape(-1, 4)
    return w


def _device_fold(w, path, grid=None, fold=None):
    g = G.create(w, path=PATHS[path])
    if grid is not None:
        G.set_plain_sdf(g, grid)
    if fold is not None:
        G.set_collision_fold(g, fold)
    G.upload(g)
    x0, y0 = np.float32(w.px).copy(), np.float32(w.py).copy()
    G.step(g)
    G.download(g)
    G.destroy(g)
    m = np.arange(0, w.n, 21)
    return np.float32(w.px)[m] - x0[m], np.float32(w.py)[m] - y0[m]


@pytest.mark.parametrize("case", ["map", "quirk"])
def test_sdf_fold_device_paths_match_the_oracle(golden, case):
    """sdf.json's queries (600 random points plus edges, corners, just-outside
    and interior points of every rect of makeMap(1200, 800, {seed: 1}); the
    best^2-quirk cases on their own rect list) and points at the cull
    boundary (radius, radius + 2, the registration margin, on both sides) go
    through the plain linear scan, the plain cull grid and both persistent
    folds: all four move the probe bit-identically, and the move is the
    oracle's n * (r - d) wherever r - d clears the dead zone by 1e-3 px."""
    g_ = golden("sdf.json")
    r = 24.0
    if case == "map":
        rects = np.asarray(po.make_map(1200, 800, 1), np.float32).astype(np.float64)
        pts = [q[:2] for q in g_["main"]]
        for rx, ry, rw, rh in rects:
            for dd in (r - 0.01, r + 0.01, r + 2.0, r + 3.0, 2 * r):
                pts += [[rx - dd, ry + rh / 2], [rx + rw + dd, ry + rh / 3], [rx + rw / 2, ry - dd],
                        [rx + rw / 4, ry + rh + dd], [rx - dd * 0.7071, ry - dd * 0.7071]]
    else:
        rects = np.asarray(g_["quirk"]["rects"], np.float64)
        pts = [q[:2] for q in g_["quirk"]["queries"]] + [[55.0, 60.0], [42.0, 15.0], [36.0, 47.0]]
    pts = np.asarray(pts, np.float64)
    pts32 = np.float32(pts).astype(np.float64)
    res = {}
    for name, path, grid, fold in (("plain-linear", "plain", 0, None), ("plain-grid", "plain", 1, None),
                                   ("persist-parallel", "persistent", None, K.KX_FOLD_PARALLEL),
                                   ("persist-owner", "persistent", None, K.KX_FOLD_OWNER)):
        say(f"[case] sdf fold {case} via {name}")
        res[name] = _device_fold(_probe_world(rects, pts32, r), path, grid, fold)
    base = res["plain-linear"]
    for name, (dx, dy) in res.items():
        assert np.array_equal(dx, base[0]) and np.array_equal(dy, base[1]), name
    checked = 0
    for i, (x, y) in enumerate(pts32):
        hit = po.sdf(np.float32(rects).astype(np.float64), x, y)
        pen = r - hit[0] if hit else -1.0
        dx, dy = float(base[0][i]), float(base[1][i])
        if pen > 2e-3 + 1e-3:
            assert abs(dx - hit[1] * pen) < 1e-3 and abs(dy - hit[2] * pen) < 1e-3, (i, x, y, hit, dx, dy)
            checked += 1
        elif pen < 2e-3 - 1e-3:
            assert dx == 0.0 and dy == 0.0, (i, x, y, hit, dx, dy)
    say(f"    {len(pts)} queries, 4 device folds bit-identical, {checked} penetrations equal to the oracle's")
    assert checked > (2 if case == "quirk" else 100)
