"""The continuous-batching server on the GPU: a served request equals the
same rollout through a handle of its own, each result is the same alone and
in a mix (bit for bit: the batch only decides which workgroups share a
launch), admission holds requests back by free HBM, and bad requests are
refused."""
import ctypes as C
import sys
import time

import numpy as np
import pytest

from kraken_3_gpu_amd import _lib as K
from kraken_3_gpu_amd import gpuworld as G
from kraken_3_gpu_amd import serve as S
from kraken_3_gpu_amd import world as W
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
KEYS = ("px", "py", "ang", "vx", "vy", "av")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _req(tag, arms, segs, ring, pairs, steps, x=420.0, y=330.0, angle=0.7):
    return dict(arms=arms, segs=segs, ring=ring, pairs=pairs, steps=steps, x=x, y=y, angle=angle,
                policy_seed=1000 + tag, noise_seed=2000 + 7 * tag, sigma=S.SIGMA, tag=tag)


def _drain(srv, n, timeout=120.0):
    out = []
    t0 = time.perf_counter()
    while len(out) < n and time.perf_counter() - t0 < timeout:
        out.extend(srv.poll(wait=0.05))
    assert len(out) == n, f"{len(out)} of {n} completed"
    return {r["tag"]: r for r in out}


def test_served_request_equals_handle_rollout():
    """Server (chunks of 50 steps, batched launch, tapes from its policy
    workers) vs one handle of the same world stepped by kx_rollout over the
    whole kx_policy_tape: identical final bodies and fitness."""
    d = _req(1, arms=5, segs=6, ring=7, pairs=6, steps=130)
    say("[case] server single request")
    with S.Server(S.default_opts(chunk_steps=50)) as srv:
        srv.submit(d, want_state=True)
        r = _drain(srv, 1)[1]
    assert r["status"] == 0 and r["n_creatures"] == 12
    b = W.Builder()
    m = S.MAP
    b.map(m["width"], m["height"], m["seed"], m["cols"], m["rows"], m["density"], m["wall"])
    for _ in range(2 * d["pairs"]):
        b.spawn(d["x"], d["y"], d["angle"], d["arms"], d["segs"], d["ring"])
    w = b.world()
    px0 = np.float32(w.px).copy()
    g = G.create(w, path=K.KX_PATH_PERSISTENT, lanes_per_creature=16, waves_per_block=1)
    G.upload(g)
    G.upload_tape(g, S.policy_tape(d, 0, d["steps"]))
    law = W.default_law(0)
    law.flags = K.KX_LAW_TAPE
    say("[case] same request through one handle")
    G.rollout(g, law, 0, d["steps"])
    G.download(g)
    G.destroy(g)
    for k in KEYS:
        assert np.array_equal(r["final_state"][k], np.float32(getattr(w, k))), k
    bpc = S.bodies_per_creature(d["arms"], d["segs"])
    assert np.array_equal(r["fitness"], np.float32(w.px)[::bpc] - px0[::bpc])
    assert np.abs(r["fitness"]).max() > 0


def test_each_result_same_alone_and_in_mix():
    """Six requests of different morphology, population and length: each run
    alone on a fresh server, then all together with staggered arrivals (so
    requests join and leave the batch mid-flight): bit-identical results."""
    reqs = [_req(1, 4, 5, 6, 8, 140), _req(2, 8, 9, 10, 20, 260, x=600, y=420), _req(3, 6, 7, 8, 3, 75),
            _req(4, 5, 9, 7, 40, 190, x=300, y=500), _req(5, 7, 6, 9, 16, 310), _req(6, 4, 8, 6, 1, 120)]
    alone = {}
    for d in reqs:
        say(f"[case] alone tag={d['tag']}")
        with S.Server(S.default_opts(chunk_steps=50)) as srv:
            srv.submit(d, want_state=True)
            alone.update(_drain(srv, 1))
    say("[case] mix")
    with S.Server(S.default_opts(chunk_steps=50)) as srv:
        for i, d in enumerate(reqs):
            srv.submit(d, want_state=True)
            time.sleep(0.004 * i)
        mixed = _drain(srv, len(reqs))
        info = srv.info()
    say(f"    mix launches={info['launches']} peak_active={info['peak_active']}")
    assert info["peak_active"] >= 2
    for d in reqs:
        a, m = alone[d["tag"]], mixed[d["tag"]]
        assert a["status"] == 0 and m["status"] == 0
        for k in KEYS:
            assert np.array_equal(a["final_state"][k], m["final_state"][k]), (d["tag"], k)
        assert np.array_equal(a["fitness"], m["fitness"])


def _oracle_request(d):
    """The request's population in the oracle: 2*pairs creatures of its
    shape at its pose on the server's map (binary32 rect values), the
    starting poses rounded to binary32 as the server's upload rounds them."""
    s = W.SOLVER
    o = po.World(s["substeps"], s["iterations"], s["h"])
    m = S.MAP
    rects = po.make_map(m["width"], m["height"], m["seed"], m["cols"], m["rows"], m["density"], m["wall"])
    o.set_map(m["width"], m["height"], np.float32(rects).astype(np.float64))
    for _ in range(2 * d["pairs"]):
        o.spawn(d["x"], d["y"], d["angle"], d["arms"], d["segs"], d["ring"])
    o.color()
    b = o.bodies()
    z = np.zeros(o.n)
    f = lambda a: np.float32(a).astype(np.float64)
    o.set_bodies(f(b["px"]), f(b["py"]), f(b["ang"]), z, z, z)
    return o


def _oracle_steps(o, d, step0, n):
    """n steps of applyControl (src/sim/core.js:183-220) on each creature's
    control-tape record from the request's policy, then the physics step."""
    tape = S.policy_tape(d, step0, n).astype(np.float64)
    A = d["arms"]
    for s in range(n):
        for ci in range(2 * d["pairs"]):
            r = tape[s, ci]
            o.apply_control(ci, r[0], r[1::3][:A], r[2::3][:A], r[3::3][:A])
        o.step()


def test_served_requests_match_the_oracle_one_step():
    """Three requests of different shape served together in one batched
    launch, one step each: every served creature within 1e-3 px of the
    oracle's applyControl of the same policy record and step."""
    reqs = [_req(11, 4, 5, 6, 4, 1, x=420, y=330, angle=0.7), _req(12, 8, 9, 10, 2, 1, x=610, y=420, angle=2.0),
            _req(13, 5, 8, 7, 3, 1, x=300, y=500, angle=4.1)]
    say("[case] served one-step requests vs oracle")
    with S.Server(S.default_opts(chunk_steps=50)) as srv:
        for d in reqs:
            srv.submit(d, want_state=True)
        res = _drain(srv, len(reqs))
    for d in reqs:
        r = res[d["tag"]]
        assert r["status"] == 0
        o = _oracle_request(d)
        _oracle_steps(o, d, 0, 1)
        b = o.bodies()
        err = max(float(np.abs(r["final_state"][k] - b[k]).max()) for k in ("px", "py"))
        say(f"    request {d['tag']} ({d['arms']}x{d['segs']}, ring {d['ring']}, {2 * d['pairs']} creatures): "
            f"max|dpos| {err:.3e} px")
        assert err <= 1e-3


def test_served_request_event_window_matches_the_oracle():
    """A served 24-step request equals a handle's rollout of the same tape
    bit for bit; that rollout's grip events (attach, release, break) over
    the 24 steps equal the oracle's exactly, and the final poses stay within
    0.05 px of the oracle's."""
    from helpers import decode_events
    d = _req(21, 6, 7, 8, 6, 24, x=150.0, y=300.0, angle=3.0)
    say("[case] served 24-step request vs handle vs oracle")
    with S.Server(S.default_opts(chunk_steps=50)) as srv:
        srv.submit(d, want_state=True)
        r = _drain(srv, 1)[d["tag"]]
    b = W.Builder()
    m = S.MAP
    b.map(m["width"], m["height"], m["seed"], m["cols"], m["rows"], m["density"], m["wall"])
    for _ in range(2 * d["pairs"]):
        b.spawn(d["x"], d["y"], d["angle"], d["arms"], d["segs"], d["ring"])
    w = b.world()
    g = G.create(w, path=K.KX_PATH_PERSISTENT, lanes_per_creature=16, waves_per_block=1)
    G.upload(g)
    G.upload_tape(g, S.policy_tape(d, 0, d["steps"]))
    law = W.default_law(0)
    law.flags = K.KX_LAW_TAPE
    ev = decode_events(G.rollout(g, law, 0, d["steps"], events=True))
    G.download(g)
    G.destroy(g)
    for k in KEYS:
        assert np.array_equal(r["final_state"][k], np.float32(getattr(w, k))), k
    o = _oracle_request(d)
    _oracle_steps(o, d, 0, d["steps"])
    ev_or = sorted(o.events())
    bo = o.bodies()
    err = max(float(np.abs(r["final_state"][k] - bo[k]).max()) for k in ("px", "py"))
    say(f"    24 steps: {len(ev_or)} oracle events, {len(ev)} served, final max|dpos| {err:.3e} px")
    assert len(ev_or) > 10 and ev == ev_or
    assert err <= 0.05


def _free_bytes():
    f, t = C.c_int64(), C.c_int64()
    K.check(K.lib().kx_mem_info(C.byref(f), C.byref(t)))
    return f.value


def test_admission_holds_back_by_free_memory():
    """With the reserve set so that free HBM covers one request's estimate
    but not two, four requests submitted together run one at a time, in
    arrival order; all complete, and each held no more memory than admission
    reserved for it."""
    d0 = _req(1, 6, 7, 8, 200, 60)
    est = K.lib().kx_request_bytes(S.make_request(d0), 50)
    f0 = _free_bytes()
    with S.Server(S.default_opts(chunk_steps=50)):
        own = f0 - _free_bytes()            # the server's own launch tables
    reserve = f0 - own - int(1.5 * est)
    say(f"[case] admission est={est} free={f0} server={own} reserve={reserve}")
    with S.Server(S.default_opts(chunk_steps=50, hbm_reserve=reserve)) as srv:
        for i in range(4):
            srv.submit(dict(d0, tag=i + 1, policy_seed=50 + i))
        res = _drain(srv, 4)
        info = srv.info()
    say(f"    peak_active={info['peak_active']} admit_waits={info['admit_waits']}")
    assert all(r["status"] == 0 for r in res.values())
    assert info["peak_active"] == 1 and info["admit_waits"] > 0
    assert all(0 < r["device_bytes"] <= est for r in res.values())
    order = sorted(res.values(), key=lambda r: r["t_admit"])
    assert [r["tag"] for r in order] == [1, 2, 3, 4]          # FIFO admission
    for a, b in zip(order, order[1:]):
        assert b["t_admit"] >= a["t_done"] - 1e-3


def test_bad_requests_are_refused():
    with S.Server(S.default_opts()) as srv:
        for bad in (dict(arms=9), dict(segs=0), dict(pairs=0), dict(steps=0), dict(x=float("nan"))):
            d = dict(_req(1, 5, 6, 7, 2, 10), **bad)
            with pytest.raises(K.KxError):
                srv.submit(d)
