"""GPU parity of both device paths against the CPU oracle.

Tolerance (binary32 device vs binary64 oracle, one step from identical
f32-rounded state): max position error <= 1e-3 px. Creatures whose discrete
events (grip attach / release / break) differ in that step are reported
separately and excluded from the position bound; the event sequences
themselves must match exactly over the first step window.
"""
import sys

import numpy as np
import pytest

from helpers import decode_events, load_state, load_state_subset, oracle_subset, oracle_twin
from kraken_3_gpu_amd import _lib as K
from kraken_3_gpu_amd import gpuworld as G
from kraken_3_gpu_amd import world as W
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

PATHS = {"plain": K.KX_PATH_PLAIN, "persistent": K.KX_PATH_PERSISTENT}
TOL_PX = 1e-3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _setup(maker, path):
    w = maker()
    o = oracle_twin(w)
    g = G.create(w, path=PATHS[path])
    G.upload(g)
    return w, o, g


def _pull_state(g):
    """Device state -> g.cpu (bodies, lambda, broken, control fields, cooldown)."""
    w = g.cpu
    G.download(g)
    ctl = G.download_controls(g)
    w.lbx[:], w.lby[:] = ctl["lbx"], ctl["lby"]
    w.compliance[:], w.target[:], w.breakForce[:] = ctl["compliance"], ctl["target"], ctl["breakForce"]
    G.download_cooldown(g)


def _oracle_events(o):
    return sorted(o.events())


def _creatures_with_event_diff(ev_a, ev_b):
    sa, sb = set(ev_a), set(ev_b)
    return {e[1] for e in sa ^ sb}


def _pos_err(w, o, creature_of_body, exclude):
    b = o.bodies()
    err = np.maximum(np.abs(w.px - b["px"]), np.abs(w.py - b["py"]))
    mask = ~np.isin(creature_of_body, list(exclude)) if exclude else np.ones(len(err), bool)
    return float(err[mask].max()) if mask.any() else 0.0, err


@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_coloring_and_layout_match_oracle(path):
    say(f"[case] coloring path={path}")
    w, o, g = _setup(lambda: W.world_uneven(24, seed=3, sort_by_size=False), path)
    c = o.cons()
    assert np.array_equal(w.type, c["type"]) and np.array_equal(w.a, c["a"]) and np.array_equal(w.b, c["b"])
    off, cnt = o.color_info()
    assert w.colorOff.tolist() == off and w.colorCnt.tolist() == cnt
    G.destroy(g)


@pytest.mark.parametrize("path", ["plain", "persistent"])
@pytest.mark.parametrize("warm", [0, 12, 40])
def test_one_step_rollout_parity(path, warm):
    """Device control law + reflex + step vs oracle, from identical state."""
    from helpers import creature_of_bodies
    seed = 2024
    w, o, g = _setup(lambda: W.world_cpu64(32, seed=1), path)
    law = W.default_law(seed)
    if warm:
        say(f"[case] warm-up rollout path={path} steps={warm}")
        G.rollout(g, law, 0, warm)
    _pull_state(g)
    load_state(o, w)
    o.set_step_counter(warm)
    o.clear_events()
    o.apply_law(seed, warm)
    o.step()
    say(f"[case] one-step rollout path={path} warm={warm}")
    ev_gpu = decode_events(G.rollout(g, law, warm, 1, events=True))
    ev_gpu = [(warm + s, c, a, k) for (s, c, a, k) in ev_gpu]
    G.download(g)
    ev_or = _oracle_events(o)
    flips = _creatures_with_event_diff(ev_or, ev_gpu)
    crb = creature_of_bodies(w)
    err, _ = _pos_err(w, o, crb, flips)
    say(f"    path={path} warm={warm} max|dpos|={err:.3e} px events={len(ev_or)} flips={sorted(flips)}")
    assert len(flips) <= 1
    assert err <= TOL_PX


@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_host_driven_step_parity(path):
    """The drop-in path: controls packed on the host (packControls) then
    kx_step, exactly the gpuTrainStep order (gputrain.js:93-100)."""
    from helpers import creature_of_bodies
    seed = 7
    w, o, g = _setup(lambda: W.world_cpu64(16, seed=2), path)
    law = W.default_law(seed)
    say(f"[case] host-driven warm-up path={path}")
    G.rollout(g, law, 0, 8)
    _pull_state(g)
    load_state(o, w)
    o.apply_law(seed, 8)                    # controls + reflex on the host (oracle)
    c = o.cons()
    w.compliance[:], w.target[:] = c["compliance"], c["target"]
    w.lbx[:], w.lby[:], w.breakForce[:] = c["lbx"], c["lby"], c["breakForce"]
    w.broken[:] = c["broken"]
    G.upload_controls(g)
    load_state(o, w)                        # identical f32-rounded controls on both sides
    o.step()
    say(f"[case] host-driven kx_step path={path}")
    G.step(g)
    G.download(g)
    err, _ = _pos_err(w, o, creature_of_bodies(w), set())
    say(f"    host-driven path={path} max|dpos|={err:.3e}")
    assert err <= TOL_PX
    assert np.array_equal(w.broken, o.cons()["broken"].astype(np.uint8))


@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_event_sequence_first_window(path):
    seed, window = 99, 24
    w, o, g = _setup(lambda: W.world_cpu64(32, seed=1), path)
    load_state(o, w)
    law = W.default_law(seed)
    for s in range(window):
        o.apply_law(seed, s)
        o.step()
    say(f"[case] event window path={path} steps={window}")
    ev_gpu = decode_events(G.rollout(g, law, 0, window, events=True))
    ev_or = _oracle_events(o)
    say(f"    events oracle={len(ev_or)} gpu={len(ev_gpu)}")
    assert len(ev_or) > 50
    assert ev_gpu == ev_or


@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_passive_rest_is_exact(path):
    """G1 (gpu.html:223-268): no drive -> bit-exact fixpoint, v stays 0."""
    w, o, g = _setup(lambda: W.world_cpu64(4, seed=11), path)
    px0, py0 = np.float32(w.px).copy(), np.float32(w.py).copy()
    say(f"[case] passive path={path}")
    G.step(g, 240)
    st = G.download(g)
    assert np.array_equal(np.float32(w.px), px0) and np.array_equal(np.float32(w.py), py0)
    assert st["maxSpeed"] == 0.0 and st["nan"] == 0


@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_determinism(path):
    runs = []
    for _ in range(2):
        w, o, g = _setup(lambda: W.world_cpu64(16, seed=5), path)
        say(f"[case] determinism path={path}")
        G.rollout(g, W.default_law(3), 0, 60)
        G.download(g)
        runs.append((w.px.copy(), w.py.copy(), w.broken.copy()))
        G.destroy(g)
    assert all(np.array_equal(a, b) for a, b in zip(*runs))


def test_graph_replay_matches_eager_launches():
    law = W.default_law(11)
    w1, _, g1 = _setup(lambda: W.world_uneven(64, seed=2), "persistent")
    w2, _, g2 = _setup(lambda: W.world_uneven(64, seed=2), "persistent")
    say("[case] graph vs eager")
    for r in range(4):
        G.rollout(g1, law, 5 * r, 5)
    G.rollout_graph(g2, law, 0, 5, 4)
    G.download(g1)
    G.download(g2)
    assert np.array_equal(w1.px, w2.px) and np.array_equal(w1.ang, w2.ang)
    assert np.array_equal(w1.broken, w2.broken)


def test_persistent_shapes_are_bit_identical():
    """Any lanes-per-creature / waves-per-workgroup shape runs a topological
    order of the same constraint dependencies: results must not depend on it.
    A lane holds at most 8 bodies' start poses in registers, so 8 lanes take
    creatures of up to 64 bodies: the uneven world (up to 82) runs the wider
    shapes and refuses 8 lanes with the reason; the uniform world (43 bodies)
    runs every shape."""
    law = W.default_law(21)
    for make, shapes in ((lambda: W.world_uneven(40, seed=6), ((64, 1), (32, 2), (16, 1), (16, 4))),
                         (lambda: W.world_uniform(64, seed=6), ((32, 1), (16, 1), (16, 4), (8, 1), (8, 2)))):
        ref = None
        for shape in shapes:
            w = make()
            g = G.create(w, path=K.KX_PATH_PERSISTENT, lanes_per_creature=shape[0], waves_per_block=shape[1])
            G.upload(g)
            say(f"[case] shape {shape}")
            ev = G.rollout(g, law, 0, 30, events=True)
            G.download(g)
            cur = (w.px.copy(), w.py.copy(), w.ang.copy(), w.broken.copy(), ev)
            if ref is None:
                ref = cur
            else:
                assert all(np.array_equal(x, y) for x, y in zip(ref, cur)), shape
            G.destroy(g)
    with pytest.raises(K.KxError, match="needs at least 9 lanes"):
        G.create(W.world_uneven(40, seed=6), path=K.KX_PATH_PERSISTENT, lanes_per_creature=8)
    # the same world still runs the plain path
    g = G.create(W.world_uneven(40, seed=6), path=K.KX_PATH_PLAIN, lanes_per_creature=8)
    G.destroy(g)


def test_mixed_lanes_per_template_are_bit_identical():
    """One launch with a different lane count per morphology template (the
    per-bucket shapes the autotuner picks) gives exactly the uniform-shape
    result, events included."""
    law = W.default_law(23)
    ref = None
    for mode in ("uniform", "mixed"):
        w = W.world_uneven(256, seed=7)
        g = G.create(w, path=K.KX_PATH_PERSISTENT, lanes_per_creature=16, waves_per_block=1)
        tinfo = G.template_info(g)
        assert len(tinfo) > 4 and sum(t["creatures"] for t in tinfo) == 256
        if mode == "mixed":
            # 8 lanes only for creatures of up to 64 bodies (8 per lane)
            want = [16 if (t % 4 == 0 and tinfo[t]["bodies"] > 64) else (8, 16, 32, 64)[t % 4]
                    for t in range(len(tinfo))]
            assert 8 in want
            G.set_template_lanes(g, want)
            assert G.persist_info(g)["mixed_lanes"] == 1
            assert [t["lanes"] for t in G.template_info(g)] == want
        G.upload(g)
        say(f"[case] per-template lanes {mode}: {len(tinfo)} templates")
        ev = G.rollout(g, law, 0, 25, events=True)
        G.download(g)
        cur = (w.px.copy(), w.py.copy(), w.ang.copy(), w.broken.copy(), ev)
        if ref is None:
            ref = cur
        else:
            assert all(np.array_equal(x, y) for x, y in zip(ref, cur))
        G.destroy(g)


def test_plain_and_persistent_agree_uneven():
    """Ragged morphologies (4-8 arms x 5-9 segments, ring 6-10) through both paths."""
    from helpers import creature_of_bodies
    law = W.default_law(5)
    wp, o, gp = _setup(lambda: W.world_uneven(48, seed=4), "plain")
    wq, _, gq = _setup(lambda: W.world_uneven(48, seed=4), "persistent")
    load_state(o, wp)
    o.apply_law(5, 0)
    o.step()
    say("[case] uneven one step plain+persistent")
    evp = decode_events(G.rollout(gp, law, 0, 1, events=True))
    evq = decode_events(G.rollout(gq, law, 0, 1, events=True))
    G.download(gp)
    G.download(gq)
    ev_or = _oracle_events(o)
    crb = creature_of_bodies(wp)
    for w, ev in ((wp, evp), (wq, evq)):
        flips = _creatures_with_event_diff(ev_or, ev)
        err, _ = _pos_err(w, o, crb, flips)
        say(f"    uneven max|dpos|={err:.3e} flips={sorted(flips)}")
        assert err <= TOL_PX and len(flips) <= 1


def _random_tape(rng, rows, ncr, kind="clamps"):
    """Tape records. "clamps": beyond every clamp (tone outside [0, 1],
    shoulder and curl past their limits, adhesion below the release and above
    the attach thresholds). "moderate": inside the limits, tone 0.3-0.7,
    adhesion 0-1 (the seeded law's ranges)."""
    t = np.zeros((rows, ncr, K.KX_TAPE_WORDS), np.float32)
    wide = kind == "clamps"
    t[:, :, 0] = rng.uniform(-0.2, 1.2, (rows, ncr)) if wide else rng.uniform(0.3, 0.7, (rows, ncr))
    for a in range(8):
        t[:, :, 1 + 3 * a] = rng.uniform(-3.0, 3.0, (rows, ncr)) if wide else rng.uniform(-1.0, 1.0, (rows, ncr))
        t[:, :, 2 + 3 * a] = rng.uniform(-2.0, 2.0, (rows, ncr)) if wide else rng.uniform(-0.8, 0.8, (rows, ncr))
        t[:, :, 3 + 3 * a] = rng.uniform(-0.3, 1.1, (rows, ncr)) if wide else rng.uniform(0.0, 1.0, (rows, ncr))
    return t


def _tape_law():
    law = W.default_law(0)
    law.flags = K.KX_LAW_TAPE
    return law


# Records past the clamps drive joints onto their limits at random tone; a
# pin or motor dead-zone comparison (2e-3 px, 2e-4 rad) that binary32 and
# binary64 decide differently then moves a body by up to the dead zone (the
# oracle's own binary32 emulation of the reference shows the same, 1.0-1.6e-3
# px on these worlds), so that case bounds positions by the dead zone plus
# arithmetic error and pins the control mapping itself bit for bit.
TOL_PX_CLAMPS = 2.5e-3


@pytest.mark.parametrize("kind", ["moderate", "clamps"])
@pytest.mark.parametrize("path", ["plain", "persistent"])
def test_tape_step_parity(path, kind):
    """KX_LAW_TAPE: applyControl of each creature's tape record (core.js:183-220)
    then one step, against the oracle's applyControl from identical state;
    4-8 arms per creature, grips attached and free. Targets, compliances and
    break forces must equal the oracle's rounded to binary32; events must
    match; positions within 1e-3 px (moderate records) or the dead-zone
    bound (records past the clamps)."""
    from helpers import creature_of_bodies
    w, o, g = _setup(lambda: W.world_uneven(24, seed=3, sort_by_size=False), path)
    say(f"[case] tape warm-up path={path} kind={kind}")
    G.rollout(g, W.default_law(11), 0, 12)
    _pull_state(g)
    load_state(o, w)
    o.set_step_counter(12)
    o.clear_events()
    tape = _random_tape(np.random.default_rng(5), 1, w.n_creatures, kind)
    for ci in range(w.n_creatures):
        arms = int(w.arm_off[ci + 1] - w.arm_off[ci])
        r = tape[0, ci].astype(np.float64)
        o.apply_control(ci, r[0], r[1::3][:arms], r[2::3][:arms], r[3::3][:arms])
    o.step()
    G.upload_tape(g, tape)
    say(f"[case] tape one step path={path} kind={kind}")
    ev_gpu = [(12 + s, c, a, k) for (s, c, a, k) in decode_events(G.rollout(g, _tape_law(), 12, 1, events=True))]
    G.download(g)
    ev_or = _oracle_events(o)
    flips = _creatures_with_event_diff(ev_or, ev_gpu)
    err, _ = _pos_err(w, o, creature_of_bodies(w), flips)
    say(f"    tape path={path} kind={kind} max|dpos|={err:.3e} px events={len(ev_or)} flips={sorted(flips)}")
    assert len(ev_or) > 0
    assert len(flips) <= 1
    assert err <= (TOL_PX if kind == "moderate" else TOL_PX_CLAMPS)
    ctl, oc = G.download_controls(g), o.cons()
    mot = w.type == K.KX_ANGLE
    for key in ("compliance", "target"):
        assert np.array_equal(ctl[key][mot], oc[key][mot].astype(np.float32)), key
    grip = (w.type == K.KX_GRIP) & (w.broken == oc["broken"])
    assert np.array_equal(ctl["breakForce"][grip], oc["breakForce"][grip].astype(np.float32))
    G.destroy(g)


def test_tape_rollout_paths_and_graph_bit_identical():
    """A 16-row tape drives 16 steps identically on the plain and persistent
    paths; a captured rollout replays rows 0..7 on every replay, exactly as
    eager rollouts over the same rows; too few rows is an error."""
    make = lambda: W.world_uneven(256, seed=9)
    tape = None
    states = []
    for path in ("plain", "persistent"):
        w = make()
        if tape is None:
            tape = _random_tape(np.random.default_rng(3), 16, w.n_creatures)
        g = G.create(w, path=PATHS[path])
        G.upload(g)
        G.upload_tape(g, tape)
        say(f"[case] tape 16 steps path={path}")
        ev = G.rollout(g, _tape_law(), 0, 16, events=True)
        G.download(g)
        states.append((w.px.copy(), w.py.copy(), w.ang.copy(), w.broken.copy(), ev))
        G.destroy(g)
    assert all(np.array_equal(a, b) for a, b in zip(*states))
    assert states[0][4].any()
    runs = []
    for mode in ("eager", "graph"):
        w = make()
        g = G.create(w, path=K.KX_PATH_PERSISTENT)
        G.upload(g)
        G.upload_tape(g, tape[:8])
        with pytest.raises(K.KxError):
            G.rollout(g, _tape_law(), 0, 9)
        say(f"[case] tape rows 0..7 twice, {mode}")
        if mode == "eager":
            G.rollout(g, _tape_law(), 0, 8)
            G.rollout(g, _tape_law(), 8, 8)
        else:
            G.rollout_graph(g, _tape_law(), 0, 8, 2)
        G.sync(g)
        G.download(g)
        runs.append((w.px.copy(), w.py.copy(), w.ang.copy(), w.broken.copy()))
        G.destroy(g)
    assert all(np.array_equal(a, b) for a, b in zip(*runs))


@pytest.mark.parametrize("case", ["uniform4096", "uneven512"])
def test_plain_and_persistent_bit_identical(case):
    """Both paths hold global binary32 state between steps, run each step
    relative to the same integer origin and solve the same constraints in
    orders that differ only between constraints sharing no body; the
    persistent path's two-phase collision pass reproduces the sequential one.
    So a multi-step rollout under the control law must agree bit for bit:
    positions, angles, velocities, constraint state and grip events."""
    if case == "uniform4096":
        make, steps, shapes = (lambda: W.world_uniform(4096, seed=1)), 24, ((32, 2),)
    else:
        make, steps, shapes = (lambda: W.world_uneven(512, seed=3)), 24, ((32, 2), (16, 1))
    law = W.default_law(77)
    wp = make()
    gp = G.create(wp, path=K.KX_PATH_PLAIN)
    G.upload(gp)
    say(f"[case] bit-identity {case} plain {steps} steps")
    evp = G.rollout(gp, law, 0, steps, events=True)
    G.download(gp)
    G.destroy(gp)
    for lanes, wpb in shapes:
        wq = make()
        gq = G.create(wq, path=K.KX_PATH_PERSISTENT, lanes_per_creature=lanes, waves_per_block=wpb)
        G.upload(gq)
        say(f"[case] bit-identity {case} persistent lanes={lanes} wpb={wpb}")
        evq = G.rollout(gq, law, 0, steps, events=True)
        G.download(gq)
        G.destroy(gq)
        for key in ("px", "py", "ang", "vx", "vy", "av", "broken", "lambda_"):
            a, b = getattr(wp, key), getattr(wq, key)
            nd = int(np.count_nonzero(a != b))
            if nd:
                say(f"    {key}: {nd} differ, max {np.max(np.abs(a.astype(np.float64) - b)):.3e}")
            assert nd == 0, key
        assert np.array_equal(evp, evq)
        say(f"    identical, {len(decode_events(evq))} events")


@pytest.mark.parametrize("case", ["uniform_walls", "stress_corridors"])
def test_collision_folds_bit_identical_to_plain(case):
    """The persistent path's two collision folds (parallel candidate list with
    an in-order re-run, and the owning lane's in-order fold) reproduce the
    plain path's sequential fold (xpbd.js:441-470) bit for bit, on worlds
    where bodies press against rects: the uniform batch run until its
    creatures reach the walls, and a corridor map where most bodies touch one.
    Every path runs the same rollout from the same initial world."""
    law = W.default_law(5)
    if case == "uniform_walls":
        make, steps = (lambda: W.world_uniform(1024, seed=2)), 160
    else:
        make, steps = (lambda: W.world_stress(2048, seed=4)), 60
    results = {}
    for name, path, fold in (("plain", K.KX_PATH_PLAIN, None), ("parallel", K.KX_PATH_PERSISTENT, K.KX_FOLD_PARALLEL),
                             ("owner", K.KX_PATH_PERSISTENT, K.KX_FOLD_OWNER)):
        w = make()
        g = G.create(w, path=path)
        if fold is not None:
            G.set_collision_fold(g, fold)
            assert G.persist_info(g)["collision_fold"] == fold
        G.upload(g)
        say(f"[case] collision folds {case}: {name}, {steps} steps")
        ev = G.rollout(g, law, 0, steps, events=True)
        G.download(g)
        G.destroy(g)
        results[name] = (w, ev)
    wp, evp = results["plain"]
    for name in ("parallel", "owner"):
        wq, evq = results[name]
        for key in ("px", "py", "ang", "vx", "vy", "av", "broken", "lambda_"):
            a, b = getattr(wp, key), getattr(wq, key)
            nd = int(np.count_nonzero(a != b))
            if nd:
                say(f"    {name} {key}: {nd} differ, max {np.max(np.abs(a.astype(np.float64) - b)):.3e}")
            assert nd == 0, (name, key)
        assert np.array_equal(evp, evq), name
    w0 = make()
    say(f"    plain == parallel == owner ({len(decode_events(evp))} events, max |dx| "
        f"{float(np.max(np.abs(wp.px - w0.px))):.1f} px)")


def test_full_size_uniform_properties_and_subset_parity():
    """4096 creatures (the uniform batch): limits, NaN, and one-step parity of
    a 64-creature subset against the oracle."""
    from helpers import creature_of_bodies
    seed = 31
    w = W.world_uniform(4096, seed=1)
    g = G.create(w, path=K.KX_PATH_PERSISTENT)
    G.upload(g)
    law = W.default_law(seed)
    say("[case] uniform 4096 rollout 60")
    G.rollout(g, law, 0, 60)
    st = G.download(g)
    assert st["nan"] == 0
    assert st["maxSpeed"] <= 360 * 1.001 and st["maxAngSpeed"] <= 18 * 1.001
    _pull_state(g)
    sub = list(range(0, 4096, 64))
    o, bidx, cidx = oracle_subset(w, sub)
    load_state_subset(o, w, sub, bidx, cidx)
    o.set_step_counter(60)
    o.clear_events()
    # the subset world numbers its creatures 0..63; the control law keys on
    # the batch index, so the oracle applies it per creature with that key
    _apply_law_keyed(o, law.seed, 60, sub)
    o.step()
    say("[case] uniform 4096 one more step")
    ev = G.rollout(g, law, 60, 1, events=True)
    G.download(g)
    ev_gpu = {(0, sub.index(c), a, k) for (_, c, a, k) in decode_events(ev) if c in sub}
    ev_or = {(0, c, a, k) for (_, c, a, k) in o.events()}
    flips = {e[1] for e in ev_gpu ^ ev_or}
    b = o.bodies()
    err = np.maximum(np.abs(w.px[bidx] - b["px"]), np.abs(w.py[bidx] - b["py"]))
    crb = np.repeat(np.arange(len(sub)), 43)
    ok = ~np.isin(crb, list(flips))
    say(f"    uniform subset max|dpos|={err[ok].max():.3e} (all {err.max():.3e}) flips={sorted(flips)}")
    assert len(flips) <= 1
    assert err[ok].max() <= TOL_PX


def test_stress_map_full_size_properties_and_subset_parity():
    """65 536 creatures on the 4096 x 4096 corridor map (1 461 rects, the map
    read from HBM, not LDS), every tip adhesive with a once-drawn break limit
    in [50, 200] N: limits and NaN over a rollout, then one step of a
    64-creature subset against the oracle under the same law."""
    from helpers import creature_of_bodies
    seed = 41
    w = W.world_stress(65536, seed=1)
    law = W.default_law(seed, grip_min=50.0, grip_max=200.0)
    law.attach_min, law.release_below = -1.0, -1.0
    law.flags = K.KX_LAW_FIXED_GRIP
    g = G.create(w, path=K.KX_PATH_PERSISTENT)
    G.upload(g)
    info = G.persist_info(g)
    say(f"[case] stress 65536 rollout 30, launch shape {info}")
    assert info["map_words"] == 0                  # this map does not fit in LDS
    ev = G.rollout(g, law, 0, 30, events=True)
    st = G.download(g)
    assert st["nan"] == 0
    assert st["maxSpeed"] <= 360 * 1.001 and st["maxAngSpeed"] <= 18 * 1.001
    attaches = int(((ev & 0o111111) != 0).sum())
    say(f"    stress: {attaches} creature-steps with attach events, grips held "
        f"{int((w.broken[w.type == K.KX_GRIP] == 0).sum())}")
    assert attaches > 0
    _pull_state(g)
    grips = w.type == K.KX_GRIP
    held = w.breakForce[grips & (w.broken == 0)]
    assert held.size and held.min() >= 50.0 and held.max() <= 200.0
    sub = list(range(0, 65536, 1024))
    o, bidx, cidx = oracle_subset(w, sub)
    load_state_subset(o, w, sub, bidx, cidx)
    o.set_step_counter(30)
    o.clear_events()
    for local, key in enumerate(sub):
        o.apply_law_ex(seed, 30, local, key, 50.0, 200.0, -1.0, -1.0, True)
    o.step()
    say("[case] stress 65536 one more step")
    ev1 = G.rollout(g, law, 30, 1, events=True)
    G.download(g)
    ev_gpu = {(0, sub.index(c), a, k) for (_, c, a, k) in decode_events(ev1) if c in sub}
    ev_or = {(0, c, a, k) for (_, c, a, k) in o.events()}
    flips = {e[1] for e in ev_gpu ^ ev_or}
    b = o.bodies()
    err = np.maximum(np.abs(w.px[bidx] - b["px"]), np.abs(w.py[bidx] - b["py"]))
    crb = np.repeat(np.arange(len(sub)), 43)
    ok = ~np.isin(crb, list(flips))
    say(f"    stress subset max|dpos|={err[ok].max():.3e} p99={np.percentile(err[ok], 99):.3e} flips={sorted(flips)}")
    assert len(flips) <= 1
    assert err[ok].max() <= TOL_PX
    G.destroy(g)


BODY_KEYS = ("px", "py", "ang", "vx", "vy", "av")
CONS_KEYS = ("broken", "lbx", "lby", "compliance", "target", "breakForce")


def _f64_spread(w, c, pre, seed, step, draws=16):
    """Largest change of the f64 oracle's one-step result for creature c when
    its binary32 input pose (px, py, ang) moves to a neighbouring binary32
    value (random signs, fixed seed): the oracle's own discontinuity at this
    state, where the dead zones (xpbd.js SLACK_POS / SLACK_ANG), joint limits
    and contacts switch a correction on or off."""
    rng = np.random.default_rng(0)

    def one(perturb):
        o, bidx, cidx = oracle_subset(w, [c])
        st = {k: pre[k][bidx].astype(np.float32).astype(np.float64) for k in BODY_KEYS}
        if perturb:
            for k in ("px", "py", "ang"):
                u = np.spacing(np.abs(st[k]).astype(np.float32)).astype(np.float64)
                st[k] = st[k] + u * rng.choice([-1.0, 1.0], size=len(u))
        o.set_bodies(*(st[k] for k in BODY_KEYS))
        o.set_cons_dynamic(broken=pre["broken"][cidx].astype(np.int32),
                           **{k: pre[k][cidx].astype(np.float32) for k in CONS_KEYS[1:]})
        o.set_cd(np.ascontiguousarray(pre["cd"][[c]]))
        o.set_step_counter(step)
        o.clear_events()
        _apply_law_keyed(o, seed, step, [c])
        o.step()
        b = o.bodies()
        return b["px"], b["py"]

    x0, y0 = one(False)
    return max(float(np.maximum(np.abs(x - x0), np.abs(y - y0)).max()) for x, y in (one(True) for _ in range(draws)))


def test_uneven_8192_autotuned_graph_subset_parity():
    """The uneven workload as benchmarked: 8192 creatures of 125 shapes, the
    autotuner's per-bucket lane counts and waves per workgroup, 50 steps per
    launch, a hipGraph-captured launch replayed; then one step of a
    64-creature subset (every shape bucket) against the oracle under the same
    law, and limits / NaN over the whole batch. A creature whose error
    exceeds the bar must be one at which the f64 oracle is itself
    discontinuous (a one-ulp change of its binary32 input moves the oracle's
    result by more than the bar, _f64_spread); it is reported as a dead-zone
    flip, counted with the event flips (at most one creature in all), and its
    error is bounded by the bar plus that spread. At step 150 of this seed one
    creature (8 arms x 5 segments, ring 10) is such a case: the device is
    1.07e-3 px from f64, the oracle's own one-ulp spread 1.07e-3 px, the
    WGSL-semantics f32 emulator 4.6e-4 px."""
    from kraken_3_gpu_amd import autotune as T
    seed = 13
    w = W.world_uneven(8192, seed=1)
    law = W.default_law(seed)
    say("[case] uneven 8192 autotune")
    g, rep = T.tune_buckets(w, law, steps=4, reps=1, passes=1, tiles=(50,))
    tile = rep["steps_per_launch"]
    lanes = [b["lanes"] for b in rep["buckets"]]
    say(f"    tuned lanes {lanes}, waves {rep['waves_per_block']}, tile {tile}, mixed "
        f"{rep['launch_shape']['mixed_lanes']}")
    say("[case] uneven 8192 graph replay")
    G.rollout_graph(g, law, 0, tile, 3)
    steps = 3 * tile
    st = G.download(g)
    assert st["nan"] == 0
    assert st["maxSpeed"] <= 360 * 1.001 and st["maxAngSpeed"] <= 18 * 1.001
    _pull_state(g)
    pre = {k: np.array(getattr(w, k)) for k in BODY_KEYS + CONS_KEYS + ("cd",)}
    sub = list(range(0, 8192, 128))
    o, bidx, cidx = oracle_subset(w, sub)
    load_state_subset(o, w, sub, bidx, cidx)
    o.set_step_counter(steps)
    o.clear_events()
    _apply_law_keyed(o, law.seed, steps, sub)
    o.step()
    say("[case] uneven 8192 one more step")
    ev = G.rollout(g, law, steps, 1, events=True)
    G.download(g)
    ev_gpu = {(0, sub.index(c), a, k) for (_, c, a, k) in decode_events(ev) if c in sub}
    ev_or = {(0, c, a, k) for (_, c, a, k) in o.events()}
    flips = {e[1] for e in ev_gpu ^ ev_or}
    b = o.bodies()
    err = np.maximum(np.abs(w.px[bidx] - b["px"]), np.abs(w.py[bidx] - b["py"]))
    crb = np.concatenate([np.full(1 + w.spawns[c][3] * w.spawns[c][4], i) for i, c in enumerate(sub)])
    ok = ~np.isin(crb, list(flips))
    shapes = {tuple(w.spawns[c][3:6]) for c in sub}
    over = sorted({int(crb[i]) for i in np.nonzero(ok & (err > TOL_PX))[0]})
    spread = {i: _f64_spread(w, sub[i], pre, law.seed, steps) for i in over}
    inside = ok & ~np.isin(crb, over)
    say(f"    uneven subset ({len(shapes)} shapes) max|dpos|={err[inside].max():.3e} event flips={sorted(flips)}")
    for i in over:
        e = float(err[crb == i].max())
        say(f"    dead-zone flip: creature {sub[i]} shape {tuple(w.spawns[sub[i]][3:6])} |dpos| {e:.3e}, "
            f"f64 one-ulp spread {spread[i]:.3e}")
        assert spread[i] > TOL_PX and e <= TOL_PX + spread[i]
    assert len(flips) + len(over) <= 1
    assert err[inside].max() <= TOL_PX
    G.destroy(g)


def _apply_law_keyed(o, seed, step, keys):
    """or_apply_law keys the PRNG by the oracle's own creature index; a
    subset needs the batch index, so remap through a seed-preserving loop."""
    lib = po.lib()
    import ctypes as C
    for local, key in enumerate(keys):
        lib.or_apply_law_one(o.p, C.c_uint32(seed), C.c_uint32(step), local, key)


@pytest.mark.parametrize("world", ["uniform", "uneven"])
def test_state_readout_matches_oracle(world):
    """kx_read_state after each of six steps: readProprio (the oracle's
    restatement of src/sim/core.js:129-175, pinned to the reference's own
    outputs) of every creature on the downloaded binary32 state, the grip
    load from |lambda| / h as _postGripBreak measures it, the acceleration
    from the previous readout's mantle velocity; then the mantle state. Bar:
    binary32 rounding of binary64 values (1e-6 relative + 1e-6)."""
    w = W.world_uniform(64, seed=3) if world == "uniform" else W.world_uneven(64, seed=5)
    g = G.create(w, path=K.KX_PATH_PERSISTENT)
    G.upload(g)
    law = W.default_law(9)
    h, S = w.solver["h"], w.solver["substeps"]
    arms_of = []
    for c in range(g.nCr):
        arms = []
        for a in range(w.arm_off[c], w.arm_off[c + 1]):
            ms = w.motors[w.motor_off[a]:w.motor_off[a + 1]]
            arms.append((int(w.a[ms[0]]), [int(w.b[m]) for m in ms], int(w.arm_grip[a])))
        arms_of.append(arms)
    prev = np.zeros((g.nCr, 2))
    worst = 0.0
    loads = 0
    for s in range(6):
        G.rollout(g, law, s, 1)
        st = G.read_state(g)
        G.download(g)
        bf = G.download_controls(g)["breakForce"]   # the reflex sets it on the device; download() leaves it
        say(f"[case] state readout {world} step {s + 1}")
        for c in range(g.nCr):
            arms = arms_of[c]
            A = len(arms)
            grips = []
            for (_, _, gk) in arms:
                f = float(np.hypot(np.float64(w.lambdaX[gk]), np.float64(w.lambdaY[gk]))) * (1.0 / h)
                grips.append((int(w.broken[gk]), f, float(bf[gk])))
            out, side = po.read_proprio(w.px, w.py, w.ang, w.vx, w.vy, w.av, arms[0][0], [a[1] for a in arms], grips,
                                        prev[c], h, S)
            m = arms[0][0]
            ax, ay = (w.vx[m] - prev[c][0]) / (h * S), (w.vy[m] - prev[c][1]) / (h * S)
            extra = [w.px[m], w.py[m], w.ang[m], w.vx[m], w.vy[m], w.av[m], side[3], side[0], side[1], side[2], ax, ay]
            want = np.array(out + extra, np.float64)
            got = st[c, :6 * A + 6 + K.KX_STATE_EXTRA].astype(np.float64)
            err = np.abs(got - want) / (1e-6 * np.abs(want) + 1e-6)
            worst = max(worst, float(err.max()))
            assert err.max() <= 1.0, (c, s, int(np.argmax(err)), got[np.argmax(err)], want[np.argmax(err)])
            loads += sum(1 for v in out[4 * A:5 * A] if 0.0 < v)
            prev[c] = (w.vx[m], w.vy[m])
    say(f"    {g.nCr} creatures x 6 steps within the bar (worst {worst:.2f} of it), {loads} loaded grips")
    G.destroy(g)


def test_register_slots_bit_identical():
    """The kernels specialised to the 6 x 7 schedule (a uniform batch of the
    12-round template at 16 lanes, one wavefront per SIMD): round slots held
    in registers, and the same fused hinge/muscle rounds with the slots read
    from LDS, against the run-time kernel (slots read from LDS each round,
    rounds unfused) and against the plain path: the same solves in the same
    order, so every state and event agrees bit for bit over a rollout that
    reaches the walls."""
    law = W.default_law(31)
    make, steps = (lambda: W.world_uniform(1024, seed=5)), 120
    runs = {}
    for name, path, reg in (("plain", K.KX_PATH_PLAIN, None), ("lds", K.KX_PATH_PERSISTENT, 0),
                            ("registers", K.KX_PATH_PERSISTENT, 1), ("fixed_lds", K.KX_PATH_PERSISTENT, 2)):
        w = make()
        g = G.create(w, path=path)
        if reg is not None:
            G.set_register_slots(g, reg)
            info = G.persist_info(g)
            assert info["register_rounds"] == (12 if reg == 1 else 0)
            assert info["fixed_rounds"] == (12 if reg == 2 else 0)
        G.upload(g)
        say(f"[case] register slots: {name}, {steps} steps")
        ev = G.rollout(g, law, 0, steps, events=True)
        G.download(g)
        ctl = G.download_controls(g)
        G.destroy(g)
        runs[name] = (w, ev, ctl)
    wp, evp, cp = runs["plain"]
    for name in ("lds", "registers", "fixed_lds"):
        wq, evq, cq = runs[name]
        for key in ("px", "py", "ang", "vx", "vy", "av", "broken", "lambda_", "lambdaX", "lambdaY"):
            assert np.array_equal(getattr(wp, key), getattr(wq, key)), (name, key)
        for key in cp:
            assert np.array_equal(cp[key], cq[key]), (name, key)
        assert np.array_equal(evp, evq), name
    say(f"    plain == lds == registers == fixed_lds ({len(decode_events(evp))} events)")


def test_fused_run_time_rounds_bit_identical_uneven():
    """Uneven morphologies (4-8 arms x 5-9 segments, ring 6-10) whose rounds
    alternate pins and motors: the run-time kernel with their hinge and
    muscle rounds fused (the template's own round count, descriptors read
    from LDS; pins on lane pairs, or on single lanes where that lets the
    schedule alternate, as at 8 lanes) against the same kernel solving every
    round on its own and against the plain path, at 8 (with 16 for the
    largest shapes), 16 and 32 lanes per creature, events included, over a
    rollout that reaches the walls."""
    law = W.default_law(41)
    steps = 120
    for lanes in (8, 16, 32):
        runs = {}
        for name, path, reg in (("plain", K.KX_PATH_PLAIN, None), ("unfused", K.KX_PATH_PERSISTENT, 0),
                                ("fused", K.KX_PATH_PERSISTENT, 1)):
            w = W.world_uneven(384, seed=11)
            g = G.create(w, path=path, lanes_per_creature=max(lanes, 16) if reg is not None else 16)
            if reg is not None:
                if lanes == 8:
                    # 8 lanes hold at most 64 bodies: the largest shapes stay at 16
                    G.set_template_lanes(g, [8 if t["bodies"] <= 64 else 16 for t in G.template_info(g)])
                G.set_register_slots(g, reg)
                info = G.persist_info(g)
                assert info["register_rounds"] == 0 and info["fixed_rounds"] == 0
                if reg == 0:
                    assert info["fused_creatures"] == 0
                else:
                    assert info["fused_creatures"] >= (384 if lanes == 32 else 100), info["fused_creatures"]
                    say(f"    {lanes} lanes: {info['fused_creatures']} of 384 creatures in fused templates")
            G.upload(g)
            say(f"[case] fused run-time rounds: {name}, {lanes} lanes, {steps} steps")
            ev = G.rollout(g, law, 0, steps, events=True)
            G.download(g)
            G.destroy(g)
            runs[name] = (w, ev)
        wp, evp = runs["plain"]
        for name in ("unfused", "fused"):
            wq, evq = runs[name]
            for key in ("px", "py", "ang", "vx", "vy", "av", "broken", "lambda_", "lambdaX", "lambdaY"):
                assert np.array_equal(getattr(wp, key), getattr(wq, key)), (lanes, name, key)
            assert np.array_equal(evp, evq), (lanes, name)
        say(f"    {lanes} lanes: plain == unfused == fused ({len(decode_events(evp))} events)")


def test_inverted_joint_limits_are_not_fused():
    """A motor whose limits are inverted (min > max) is outside what the
    fused motor solve computes (its limit excess as two one-sided terms), so
    its creature's template must take the one-round-at-a-time solve while the
    other creatures keep their fused rounds: the launch reports that split,
    and every state and event agrees bit for bit with the plain path."""
    law = W.default_law(43)
    steps = 120
    runs = {}
    for name, path in (("plain", K.KX_PATH_PLAIN), ("persistent", K.KX_PATH_PERSISTENT)):
        w = W.world_uniform(256, seed=3)
        k = int(w.motors[w.motor_off[0] + 2])             # a joint motor of creature 0
        w.minA[k], w.maxA[k] = 0.05, -0.05
        g = G.create(w, path=path)
        if path == K.KX_PATH_PERSISTENT:
            info = G.persist_info(g)
            assert info["register_rounds"] == 0 and info["fixed_rounds"] == 0, info
            assert info["fused_creatures"] == 255, info
        G.upload(g)
        say(f"[case] inverted limits: {name}, {steps} steps")
        ev = G.rollout(g, law, 0, steps, events=True)
        G.download(g)
        G.destroy(g)
        runs[name] = (w, ev)
    wp, evp = runs["plain"]
    wq, evq = runs["persistent"]
    for key in ("px", "py", "ang", "vx", "vy", "av", "broken", "lambda_", "lambdaX", "lambdaY"):
        assert np.array_equal(getattr(wp, key), getattr(wq, key)), key
    assert np.array_equal(evp, evq)
    say(f"    plain == persistent with one template unfused ({len(decode_events(evp))} events)")


def test_folds_bit_identical_over_a_long_rollout():
    """Every persistent variant (register-held or LDS round slots, parallel or
    owner collision fold) against the plain path over 1250 steps of the
    4096-creature uniform batch, as its creatures pile against the rects. A
    contact whose penetration sits near the 2e-3 px dead zone turns a
    one-ulp difference in a query into a different correction, so this is
    where an inlining context that rounds differently shows (the owner fold
    once parted from the plain path here, at step ~1150, before the build
    fixed where multiply-adds fuse)."""
    law = W.default_law(2024)
    steps = 1250
    runs = {}
    for name, path, reg, fold in (("plain", K.KX_PATH_PLAIN, None, None),
                                  ("reg_parallel", K.KX_PATH_PERSISTENT, 1, K.KX_FOLD_PARALLEL),
                                  ("reg_owner", K.KX_PATH_PERSISTENT, 1, K.KX_FOLD_OWNER),
                                  ("lds_owner", K.KX_PATH_PERSISTENT, 0, K.KX_FOLD_OWNER),
                                  ("fixed_lds_owner", K.KX_PATH_PERSISTENT, 2, K.KX_FOLD_OWNER)):
        w = W.world_uniform(4096, seed=1)
        g = G.create(w, path=path)
        if reg is not None:
            G.set_register_slots(g, reg)
            G.set_collision_fold(g, fold)
        G.upload(g)
        say(f"[case] long rollout: {name}, {steps} steps")
        ev = G.rollout(g, law, 0, steps, events=True)
        G.download(g)
        G.destroy(g)
        runs[name] = (w, ev)
    wp, evp = runs["plain"]
    for name in ("reg_parallel", "reg_owner", "lds_owner", "fixed_lds_owner"):
        wq, evq = runs[name]
        for key in ("px", "py", "ang", "vx", "vy", "av", "broken", "lambda_", "lambdaX", "lambdaY"):
            a, b = getattr(wp, key), getattr(wq, key)
            nd = int(np.count_nonzero(a != b))
            if nd:
                say(f"    {name} {key}: {nd} differ, max {np.max(np.abs(a.astype(np.float64) - b)):.3e}")
            assert nd == 0, (name, key)
        assert np.array_equal(evp, evq), name
    say(f"    plain == every persistent variant over {steps} steps ({len(decode_events(evp))} events)")
