"""The generalised creature (arms x segments, mantle ring) in the oracle and in
the product's world builder.

Two kinds of evidence pin it:

* tests/golden/morph.json: the reference's own spawnKraken run with
  BODY.arms / BODY.segsPerArm set to each shape (src/body/kraken.js:64-148
  reads both at call time; the mantle keeps 8 collision points there), its
  greedy colouring, a 24-step trajectory under the seeded control law with the
  grip events (f64, and the f32 emulator tools/f32solver.mjs for two shapes)
  and readProprio (src/sim/core.js:129-175) along it;
* closed forms over every shape of the uneven workload (arms 4-8 x segments
  5-9 x mantle ring 6-10, 125 shapes): counts, taper masses and inertias, ring
  points, zero spawn violation (tools/harness.js:118-162), motors at rest,
  2*arms colours in body-disjoint classes, exact passive rest
  (tools/harness.js:46-57) and readProprio of a straight-armed creature.
"""
import math

import numpy as np
import pytest

from kraken_3_gpu_amd import world as W
from oracle import pyoracle as po

SEG_LEN, SHOULDER_R, MANTLE_R = 13.5, 24.0, 18.0
SHOULDER_LIMIT, JOINT_LIMIT = math.pi * 0.62, math.pi * 0.42
ARM_REACH = SHOULDER_R + 7 * SEG_LEN
SHAPES = [(a, s, r) for a in range(4, 9) for s in range(5, 10) for r in range(6, 11)]
EV = {"attach": 0, "release": 1, "break": 2}


def _angle_delta(a, b):
    """angleDelta(a, b) of src/physics/xpbd.js:128-143 (wrap of a - b)."""
    d = (a - b) % (2 * math.pi)
    if d > math.pi:
        d -= 2 * math.pi
    elif d <= -math.pi:
        d += 2 * math.pi
    return d


# ── the reference's own shapes (morph.json) ──

def _morph_world(case, rects):
    o = po.World(case["S"], case["I"], case["h"])
    o.set_map(1200, 800, rects)
    for x, y, a in case["poses"]:
        o.spawn(x, y, a, case["arms"], case["segs"], 8, 0.28)
    o.color()
    if case["mode"] == "f32":
        o.f32ify()
    return o


def _cases(golden, mode=None):
    return [c for c in golden("morph.json")["cases"] if mode is None or c["mode"] == mode]


def test_reference_shapes_spawn_and_colouring(golden):
    """Oracle and product builder against the reference's spawnKraken and
    ensureColored for eight arm/segment shapes."""
    rects = po.make_map(1200, 800, 1)
    for case in _cases(golden, "f64"):
        A, S = case["arms"], case["segs"]
        cr = case["creation"]
        o = po.World(case["S"], case["I"], case["h"])
        o.set_map(1200, 800, rects)
        b = W.Builder()
        b.map(1200, 800, 1)
        for x, y, a in case["poses"]:
            o.spawn(x, y, a, A, S, 8, 0.28)
            b.spawn(x, y, a, A, S, 8)
        w = b.world()
        ob = o.bodies()
        assert o.n == w.n == cr["n"] == len(case["poses"]) * (1 + A * S)
        for key in ("px", "py", "ang"):
            np.testing.assert_allclose(ob[key], cr[key], rtol=2e-16, atol=2e-13)
            np.testing.assert_allclose(getattr(w, key), cr[key], rtol=2e-16, atol=2e-13)
        for key in ("im", "ii"):
            np.testing.assert_allclose(ob[key], cr[key], rtol=1e-15)
            assert np.array_equal(getattr(w, key), np.asarray(cr[key], np.float32)), key
        rows = np.array([[r[0], r[1], r[2], r[11], r[12]] for r in cr["cons"]])
        oc = o.cons()
        for j, key in enumerate(("type", "a", "b", "owner", "broken")):
            assert oc[key].tolist() == rows[:, j].astype(int).tolist(), key
        for key, mine in (("type", w.type), ("a", w.a), ("b", w.b), ("owner", w.owner)):
            assert mine.tolist() == oc[key].tolist(), key
        for j, key in ((3, "lax"), (4, "lay"), (5, "lbx"), (6, "lby"), (7, "compliance"), (8, "rest")):
            np.testing.assert_allclose(oc[key], [r[j] for r in cr["cons"]], rtol=1e-15, atol=1e-13)
        pts = np.array(cr["pts"])
        op = o.points()
        assert op["body"].tolist() == pts[:, 0].astype(int).tolist() == w.p_body.tolist()
        np.testing.assert_allclose(np.c_[op["lx"], op["ly"], op["r"]], pts[:, 1:], rtol=1e-15, atol=1e-14)
        # colouring: the oracle's ensureColored and the product's colouring
        o.color()
        off, cnt = o.color_info()
        col = case["colouring"]
        assert off == col["colorOff"] and cnt == col["colorCnt"], (A, S)
        assert o.cons()["id"].tolist() == col["sorted"]
        assert o.cons()["color"].tolist() == col["colorOf"]
        from test_abi_host import _color
        pc, order = _color(w)
        assert order.tolist() == col["sorted"] and pc[order].tolist() == col["colorOf"], (A, S)


@pytest.mark.parametrize("mode", ["f64", "f32"])
def test_reference_shapes_trajectories_and_events(golden, mode):
    """24 steps of the seeded control law on each shape: states against the
    reference's frames (f64 within the libm-vs-V8 last-bit band; the f32
    emulator bit for bit) and the grip event sequence exactly."""
    rects = po.make_map(1200, 800, 1)
    for case in _cases(golden, mode):
        A = case["arms"]
        o = _morph_world(case, rects)
        frames = {f["step"]: f["state"] for f in case["frames"]}
        exact = mode == "f32"
        for s in range(max(frames)):
            o.apply_law(case["seed"], s)
            o.step()
            if s + 1 not in frames:
                continue
            g = frames[s + 1]
            b = o.bodies()
            for key in ("px", "py", "ang", "vx", "vy", "av"):
                if exact:
                    assert b[key].tolist() == g[key], (A, case["segs"], s + 1, key)
                else:
                    tol = 1e-8 if key in ("px", "py", "ang") else 1e-6
                    assert np.abs(b[key] - np.array(g[key])).max() < tol, (A, case["segs"], s + 1, key)
            c = o.cons()
            cd = o.cd()
            for ci, arms in enumerate(g["grips"]):
                assert len(arms) == A
                for a, gr in enumerate(arms):
                    k = o.grip_pos(ci, a)
                    assert int(c["broken"][k]) == gr[0]
                    bf = c["breakForce"][k]
                    assert (None if math.isinf(bf) else bf) == gr[3]
                    assert cd[ci][a] == g["cd"][ci][a]
        ev = sorted(o.events())
        gev = sorted((e[0], e[1], e[2], EV[e[3]]) for e in case["events"])
        assert len(gev) > 50 and ev == gev, (A, case["segs"])


def test_reference_shapes_read_proprio(golden):
    """readProprio for 4-8 arms: 6A+6 inputs equal to the reference's along
    the trajectories (the acceleration input following the previous call)."""
    worst, n = 0.0, 0
    for case in _cases(golden, "f64"):
        A = case["arms"]
        for fr in case["frames"]:
            st = fr["state"]
            for ci, cr in enumerate(case["creatures"]):
                grips = [(gr[0], gr[6], gr[3]) for gr in st["grips"][ci]]
                out, side = po.read_proprio(st["px"], st["py"], st["ang"], st["vx"], st["vy"], st["av"],
                                            cr["mantle"], cr["arms"], grips, fr["prev"][ci], case["h"], case["S"])
                ref = fr["proprio"][ci]
                assert len(out) == len(ref) == 6 * A + 6
                worst = max(worst, max(abs(a - b) for a, b in zip(out, ref)))
                me, mc, anc, spd = fr["side"][ci]
                assert abs(side[0] - me) < 1e-12 and abs(side[1] - mc) < 1e-12 and side[2] == anc
                n += 1
    assert n >= 8 * 3 * 3 and worst < 1e-12


# ── closed forms over all 125 shapes ──

def _pin_violation(b, c, k):
    A, B = int(c["a"][k]), int(c["b"][k])
    ca, sa = math.cos(b["ang"][A]), math.sin(b["ang"][A])
    gax = b["px"][A] + c["lax"][k] * ca - c["lay"][k] * sa
    gay = b["py"][A] + c["lax"][k] * sa + c["lay"][k] * ca
    if B >= 0:
        cb, sb = math.cos(b["ang"][B]), math.sin(b["ang"][B])
        gbx = b["px"][B] + c["lbx"][k] * cb - c["lby"][k] * sb
        gby = b["py"][B] + c["lbx"][k] * sb + c["lby"][k] * cb
    else:
        gbx, gby = c["lbx"][k], c["lby"][k]
    return math.hypot(gax - gbx, gay - gby)


def _shape_checks(A, S, R, b, c, p, first, x, y, angle):
    """Counts, masses, ring points, zero violation and motors at rest of one
    creature whose bodies start at `first`; b/c/p are dicts of arrays."""
    nb = 1 + A * S
    mine = (c["a"] >= first) & (c["a"] < first + nb)
    ks = np.nonzero(mine)[0]
    assert len(ks) == A * (2 * S + 1)
    pts = np.nonzero((p["body"] >= first) & (p["body"] < first + nb))[0]
    assert len(pts) == R + 2 * A * S
    # mantle: m = 20, I = 20*18^2/2*2.5 (kraken.js:70)
    assert b["px"][first] == x and b["py"][first] == y and b["ang"][first] == angle
    assert abs(b["im"][first] - 1 / 20) < 1e-7 and abs(b["ii"][first] - 1 / 8100) < 1e-10
    # taper ends: segment 0 has m 3.2, thickness 11.9; the tip 1.8, 5.9 (kraken.js:84-90)
    for arm in range(A):
        s0, tip = first + 1 + arm * S, first + arm * S + S
        for body, m, t in ((s0, 3.2, 11.9), (tip, 1.8, 5.9)):
            assert abs(b["im"][body] * m - 1) < 1e-6
            assert abs(b["ii"][body] * m * (SEG_LEN ** 2 + t * t) / 12 - 1) < 1e-6
        # each segment the next one's shorter-and-lighter neighbour
        ms = [1 / b["im"][first + 1 + arm * S + i] for i in range(S)]
        assert all(ms[i] > ms[i + 1] for i in range(S - 1))
    # mantle ring: R points on radius 0.7*18 at 2*pi*i/R, radius 0.45*18
    ring = [j for j in pts if p["body"][j] == first]
    assert len(ring) == R
    for i, j in enumerate(ring):
        a = 2 * math.pi * i / R
        assert abs(p["lx"][j] - 12.6 * math.cos(a)) < 1e-5 and abs(p["ly"][j] - 12.6 * math.sin(a)) < 1e-5
        assert abs(p["r"][j] - 8.1) < 1e-5
    worst_pin, worst_motor, n_pin, n_motor = 0.0, 0.0, 0, 0
    for k in ks:
        t = int(c["type"][k])
        if t == po.ANGLE:
            A_, B_ = int(c["a"][k]), int(c["b"][k])
            rel = _angle_delta(b["ang"][B_], b["ang"][A_])
            dev = _angle_delta(rel, c["rest"][k])
            worst_motor = max(worst_motor, abs(dev - c["target"][k]))
            n_motor += 1
        elif not c["broken"][k]:
            worst_pin = max(worst_pin, _pin_violation(b, c, k))
            n_pin += 1
        else:
            assert t == po.GRIP
    assert n_pin == A * S and n_motor == A * S
    return worst_pin, worst_motor


def _oracle_shape_world(shapes, x=600.0, y=400.0):
    """All shapes in one oracle world at one spot of a border-only
    1200x800 map (creatures never interact)."""
    o = po.World()
    o.set_map(1200, 800, po.make_map(1200, 800, 1, 1, 1))
    for i, (A, S, R) in enumerate(shapes):
        o.spawn(x, y, 0.05 * i, A, S, R)
    return o


def test_all_shapes_closed_forms_oracle():
    o = _oracle_shape_world(SHAPES)
    b, c, p = o.bodies(), o.cons(), o.points()
    first = 0
    for i, (A, S, R) in enumerate(SHAPES):
        info = o.creature(i)
        assert (info["arms"], info["segs"], info["ring"], info["first_body"]) == (A, S, R, first)
        pin, motor = _shape_checks(A, S, R, b, c, p, first, 600.0, 400.0, 0.05 * i)
        assert pin < 1e-9 and motor < 1e-9, (A, S, R, pin, motor)
        first += 1 + A * S
    assert first == o.n


def test_all_shapes_closed_forms_product_builder():
    """The same checks on the product's C++ world builder (f32 statics)."""
    b = W.Builder()
    b.map(1200, 800, 1, 1, 1)
    for i, (A, S, R) in enumerate(SHAPES):
        b.spawn(600.0, 400.0, 0.05 * i, A, S, R)
    w = b.world()
    bd = dict(px=w.px, py=w.py, ang=w.ang, im=w.im.astype(np.float64), ii=w.ii.astype(np.float64))
    cd = dict(type=w.type, a=w.a, b=w.b, lax=w.lax, lay=w.lay, lbx=w.lbx, lby=w.lby, rest=w.rest,
              target=w.target, broken=w.broken)
    pd = dict(body=w.p_body, lx=w.p_lx, ly=w.p_ly, r=w.p_r)
    first = 0
    for i, (A, S, R) in enumerate(SHAPES):
        pin, motor = _shape_checks(A, S, R, bd, cd, pd, first, 600.0, 400.0, 0.05 * i)
        # anchors and rests are binary32 statics in the product
        assert pin < 1e-5 and motor < 1e-6, (A, S, R, pin, motor)
        first += 1 + A * S


def _colour_classes(o):
    off, cnt = o.color_info()
    c = o.cons()
    return off, cnt, c


@pytest.mark.parametrize("A", [4, 5, 6, 7, 8])
def test_all_shapes_colour_count_and_disjoint_classes(A):
    """The mantle carries 2A constraints (shoulder pin + motor per arm), so
    greedy colouring needs at least 2A colours; the arm chains alternate
    through four, so max(4, 2A) = 2A colours suffice; every class is
    body-disjoint."""
    for S in range(5, 10):
        for R in (6, 10):
            o = po.World()
            o.spawn(600.0, 400.0, 0.3, A, S, R)
            o.spawn(900.0, 300.0, 1.3, A, S, R)
            o.color()
            off, cnt, c = _colour_classes(o)
            used = [k for k in range(16) if cnt[k]]
            assert used == list(range(max(4, 2 * A))), (A, S, R, cnt)
            assert sum(cnt) == 2 * A * (2 * S + 1)
            for col in used:
                bodies = []
                for k in range(off[col], off[col] + cnt[col]):
                    assert c["color"][k] == col
                    bodies.append(int(c["a"][k]))
                    if c["b"][k] >= 0:
                        bodies.append(int(c["b"][k]))
                assert len(bodies) == len(set(bodies)), (A, S, R, col)


def test_all_shapes_passive_rest_is_exact():
    """harness.js 'passive' for every shape: no actuation, no contact, 240
    steps at S=4 x I=8 leave every body bit-identical and every velocity 0."""
    o = _oracle_shape_world(SHAPES)
    o.color()
    b0 = o.bodies()
    for _ in range(240):
        o.step()
    b1 = o.bodies()
    for key in ("px", "py", "ang"):
        assert np.array_equal(b0[key], b1[key]), key
    for key in ("vx", "vy", "av"):
        assert not np.any(b1[key]), key
    st = o.stats()
    assert st["maxSpeed"] == 0.0 and st["maxAngSpeed"] == 0.0 and st["nan"] == 0


@pytest.mark.parametrize("A,S", [(4, 5), (5, 9), (6, 7), (8, 6), (8, 9)])
def test_read_proprio_straight_arms_closed_form(A, S):
    """A creature spawned with curl 0 at heading th, at rest, no grip: the
    arm-a shoulder input is wrap(2*pi*a/A)/(0.62*pi) clamped to +-1.2, curls
    are 0, the tip sits (24 + 13.5*S) px out along 2*pi*a/A in the body frame,
    the extension is (13.5*S - 6.75)/118.5 (capped at 1), loads, anchors,
    accelerations, speed and spin are 0 (src/sim/core.js:129-175)."""
    th = 0.7
    o = po.World()
    o.spawn(500.0, 400.0, th, A, S, 8, 0.0)
    b = o.bodies()
    segs = [[1 + a * S + i for i in range(S)] for a in range(A)]
    grips = [(1, 0.0, 0.0)] * A
    out, side = po.read_proprio(b["px"], b["py"], b["ang"], b["vx"], b["vy"], b["av"], 0, segs, grips,
                                (0.0, 0.0), 1 / 240, 4)
    assert len(out) == 6 * A + 6
    cl = lambda v: max(-1.2, min(1.2, v))
    reach = SHOULDER_R + SEG_LEN * S
    want = [cl(_angle_delta(2 * math.pi * a / A, 0) / SHOULDER_LIMIT) for a in range(A)]
    want += [0.0] * A
    for a in range(A):
        ra = 2 * math.pi * a / A
        want += [cl(reach * math.cos(ra) / ARM_REACH), cl(reach * math.sin(ra) / ARM_REACH)]
    ext = min(1.0, (SEG_LEN * S - SEG_LEN / 2) / ARM_REACH)
    want += [0.0] * A + [0.0] * A + [0.0, 0.0, 0.0, ext, 0.0, 0.0]
    np.testing.assert_allclose(out, want, atol=1e-12)
    assert abs(side[0] - ext) < 1e-12 and abs(side[1]) < 1e-12 and side[2] == 0


# ── spawn draws and in-place respawn (src/sim/train.js:103-122, 188-229) ──

def test_crn_spawns_match_reference(golden):
    """crnSpawns from the trainer's stream: every slot's x, y, heading and
    the number of stream calls, pair-shared and not, with and without a map."""
    for case in golden("respawn.json")["crn"]:
        o = po.World()
        if case["opts"] is not None:
            op = case["opts"]
            rects = po.make_map(case["width"], case["height"], op["seed"], op.get("corridorCols", 4),
                                op.get("corridorRows", 3))
            o.set_map(case["width"], case["height"], rects)
        st = [case["s0"]]
        sp, calls = o.crn_spawns(case["pop"], st, case["pairShared"], has_map=case["opts"] is not None)
        assert sp.tolist() == case["spawns"], case["seed"]
        if case["calls"] is not None:
            assert calls == case["calls"]
        if case["pairShared"]:
            assert np.array_equal(sp[0::2][: len(sp[1::2])], sp[1::2])


def _creature_of(o, bodies):
    """Creature index of every body index in `bodies`."""
    first = [o.creature(i)["first_body"] for i in range(o.counts()[3])]
    return np.searchsorted(first, bodies, side="right") - 1


@pytest.mark.parametrize("which", [0, 1])
def test_respawn_matches_reference(golden, which):
    """respawnCreature on a world 40 steps into the seeded law (grips held,
    joints driven, bodies moving): the respawned creatures' full state, the
    untouched creature, multipliers and cooldowns, then 8 more steps."""
    r = golden("respawn.json")["respawn"][which]
    A, S = r["arms"], r["segs"]
    o = po.World(4, 8, 1 / 240)
    o.set_map(1200, 800, np.array(r["mapRects"], dtype=np.float64))
    for x, y, a in r["spawns"]:
        o.spawn(x, y, a, A, S, 8, 0.28)
    o.color()
    assert o.cons()["color"].tolist() == r["colorOf"]

    def check(ref, tol, cd=None):
        b, p = o.bodies(), o.pred()
        for key, got in (("px", b["px"]), ("py", b["py"]), ("ang", b["ang"]), ("vx", b["vx"]), ("vy", b["vy"]),
                         ("av", b["av"]), ("ppx", p["ppx"]), ("ppy", p["ppy"]), ("angP", p["angP"])):
            t = tol if key in ("px", "py", "ang", "ppx", "ppy", "angP") else 100 * tol
            assert np.abs(got - np.array(ref[key])).max() <= t, key
        c = o.cons()
        g = np.array([[v if v is not None else np.inf for v in row] for row in ref["cons"]])
        assert c["broken"].tolist() == g[:, 1].astype(int).tolist()
        for j, key in ((0, "target"), (6, "compliance"), (9, "breakForce")):
            assert np.allclose(c[key], g[:, j], rtol=1e-12, atol=100 * tol), key
        grip = c["type"] == po.GRIP
        for j, key in ((3, "lambdaX"), (4, "lambdaY")):
            assert np.abs(c[key][grip] - g[grip, j]).max() <= 1e4 * tol, key
        if cd is not None:
            assert np.array_equal(o.cd()[:, :A], np.array(cd)), "cooldowns"

    for s in range(40):
        o.apply_law(r["law"], s)
        o.step()
    check(r["before"], 1e-8, r["cdBefore"])
    # continue from the reference's own state, so the untouched creature and
    # the respawned poses compare to the last bit of libm vs V8
    o.set_bodies(*(np.array(r["before"][k]) for k in ("px", "py", "ang", "vx", "vy", "av")))
    o.set_pred(*(np.array(r["before"][k]) for k in ("ppx", "ppy", "angP")))
    for ci, x, y, a in r["respawned"]:
        o.respawn(ci, x, y, a)
    b1, p1 = o.bodies(), o.pred()
    for key in ("px", "py", "ang", "vx", "vy", "av"):
        assert np.abs(b1[key] - np.array(r["after"][key])).max() <= 1e-12, key
    for key in ("ppx", "ppy", "angP"):
        assert np.abs(p1[key] - np.array(r["after"][key])).max() <= 1e-12, key
    c = o.cons()
    g = np.array([[v if v is not None else np.inf for v in row] for row in r["after"]["cons"]])
    assert c["broken"].tolist() == g[:, 1].astype(int).tolist()
    assert c["target"].tolist() == g[:, 0].tolist()
    grip = c["type"] == po.GRIP
    released = grip & np.isin(_creature_of(o, c["a"]), [ci for ci, *_ in r["respawned"]])
    assert released.sum() == A * len(r["respawned"])
    assert not np.any(c["force"][released]) and not np.any(c["lambdaX"][released]) and not np.any(c["lambdaY"][released])
    assert np.all(c["broken"][released] == 1) and not np.any(g[released, 2])
    assert np.array_equal(o.cd()[:, :A], np.array(r["cdAfter"]))
    assert r["cdAfter"] == r["cdBefore"] and r["prevAfter"] == [[0, 0]] * 3
    for s in range(40, 48):
        o.apply_law(r["law"], s)
        o.step()
    check(r["later"], 1e-7, r["laterCd"])


def test_product_crn_spawns_match_reference(golden):
    """The product builder's crnSpawns against the reference's draws."""
    for case in golden("respawn.json")["crn"]:
        b = W.Builder()
        if case["opts"] is not None:
            op = case["opts"]
            b.map(case["width"], case["height"], op["seed"], op.get("corridorCols", 4), op.get("corridorRows", 3))
        st = [case["s0"]]
        sp = b.crn_spawns(case["pop"], st, case["pairShared"])
        assert sp.tolist() == case["spawns"], case["seed"]
    w, sp = W.world_crn(8, seed=1)
    assert w.n_creatures == 8 and np.array_equal(sp[0::2], sp[1::2])
    assert np.array_equal(w.px[[0, 43]], sp[:2, 0])
