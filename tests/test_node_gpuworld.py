"""The Node.js drop-in (node/gpuworld.js over the N-API addon node/kxpbd.node):
the reference's pure helpers byte-for-byte against the reference's own
outputs (harness `pack-controls`, `mirror-fields`, `coll-grouping`), the addon's
exports, and on a GPU the full create/upload/uploadControls/step/download loop
and a device-law rollout, bit-identical to the same calls through the C ABI
from Python."""
import json
import math
import os
import shutil
import subprocess


import numpy as np
import pytest

from helpers import creatures_to_js, world_to_js
from test_abi_host import _sim_world
from kraken_3_gpu_amd import _lib as K
from kraken_3_gpu_amd import gpuworld as G
from kraken_3_gpu_amd import world as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "node_driver.js")
ADDON = os.path.join(ROOT, "node", "kxpbd.node")
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None, reason="node is not installed")


def run_node(mode, job, tmp_path, timeout=120):
    p = tmp_path / f"job_{mode}.json"
    p.write_text(json.dumps(job))
    print(f"[case] node driver {mode}", flush=True)
    r = subprocess.run([NODE, DRIVER, mode, str(p)], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


def js_floats(xs):
    return np.array([float(x) if not isinstance(x, str) else float(x.replace("Infinity", "inf")) for x in xs])


@needs_node
def test_addon_loads_and_exports_the_surface():
    if not os.path.exists(ADDON):
        pytest.skip("node/kxpbd.node not built (run __graft_entry__.build())")
    code = ("const k=require(process.argv[1]);const g=require(process.argv[2]);"
            "console.log(JSON.stringify({k:Object.keys(k),g:Object.keys(g),abi:k.abiVersion(),avail:g.isAvailable()}))")
    r = subprocess.run([NODE, "-e", code, ADDON, os.path.join(ROOT, "node", "gpuworld.js")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert d["abi"] == 5 and d["avail"] is True
    for name in ("create", "destroy", "upload", "packControls", "step", "download", "rollout", "groupPoints",
                 "readState", "policyCreate", "policyWeights", "policyGenomes", "policyForward", "esCreate",
                 "esGeneration", "esUpdate", "esGet", "brainDestroy"):
        assert name in d["k"]
    for name in ("create", "upload", "uploadControls", "packControls", "step", "substep", "download", "mirrorInto",
                 "groupCollPoints", "destroy", "GPU_STRIDES", "isAvailable", "readState", "respawn"):
        assert name in d["g"]


@needs_node
def test_pure_helpers_match_reference_bytes(golden, tmp_path):
    g = golden("gpuworld.json")
    w = _sim_world(33, 2)
    assert w.type.tolist() == g["types"]
    for k in range(w.nc):
        w.compliance[k] = 0.123 + k * 1e-4
        w.target[k] = 0.01 * k
        w.breakForce[k] = math.inf if k % 3 == 0 else (5 + k if k % 3 == 1 else 3e38)
        w.broken[k] = 1 if k % 2 == 0 else 0
        w.lbx[k] = 11 + k * 1.1
        w.lby[k] = -22 - k * 0.7
    job = dict(world=world_to_js(w), bodyIn=g["bodyIn"], consIn=g["consIn"], nB=g["nB"],
               groupBodies=g["group"]["bodies"], shuffledBodies=g["group"]["shuffledBodies"])
    job["world"]["h"] = g["h"]
    r = run_node("pure", job, tmp_path)
    assert r["packed"] == g["packed"]
    m = g["mirrored"]
    for key in ("px", "py"):
        ref = [float("nan") if v == "NaN" else v for v in m[key]]
        np.testing.assert_array_equal(js_floats(r["mirrored"][key]), ref)
    assert js_floats(r["mirrored"]["lambda"]).tolist() == m["lambda"]
    assert r["mirrored"]["broken"] == [int(x) for x in m["broken"]]
    grip = w.type == K.KX_GRIP
    assert js_floats(r["mirrored"]["force"])[grip].tolist() == np.array(m["force"])[grip].tolist()
    assert float(r["mirrored"]["stats"]["maxSpeed"]) == m["stats"]["maxSpeed"]
    assert float(r["mirrored"]["stats"]["maxAngSpeed"]) == m["stats"]["maxAngSpeed"]
    assert r["mirrored"]["stats"]["nan"] == m["stats"]["nan"]
    gg = g["group"]
    assert r["group"]["order"] == gg["order"] and r["group"]["range"] == gg["range"]
    assert r["group"]["shuffledOrder"] == gg["shuffledOrder"] and r["group"]["shuffledRange"] == gg["shuffledRange"]
    assert r["groupThrows"] is True


def _targets(steps, ncr, seed=9):
    rng = np.random.default_rng(seed)
    return [[float(v) for v in rng.uniform(-0.8, 0.8, size=max(ncr, 1))] for _ in range(steps)]


@pytest.mark.gpu
@needs_node
def test_node_loop_and_rollout_match_c_abi(tmp_path):
    """Host-driven steps (uploadControls + step per step, then download) and a
    device-law rollout through Node give exactly what the same calls give
    through the C ABI from Python."""
    assert os.path.exists(ADDON), "node/kxpbd.node must be built"
    steps, rsteps = 6, 20
    mk = lambda: W.world_cpu64(8, seed=3)
    w0 = mk()
    law = W.default_law(2024)
    lawd = dict(seed=int(law.seed), tone=law.tone, shoulderLimit=law.shoulder_limit, jointLimit=law.joint_limit,
                stiffMin=law.stiff_min, stiffMax=law.stiff_max, gripMin=law.grip_min, gripMax=law.grip_max,
                slipMax=law.slip_max, cooldown=law.cooldown, attachMin=law.attach_min,
                releaseBelow=law.release_below)
    tg = _targets(steps, w0.n_creatures)
    stride = 6 * int(np.diff(np.asarray(w0.arm_off)).max()) + 6 + K.KX_STATE_EXTRA
    resp = dict(warm=12, creatures=[0, 5, 2], poses=[300.25, 410.5, 1.0, 820.0, 200.0, 5.5, 640.0, 480.0, 3.0])
    job = dict(world=world_to_js(w0), creatures=creatures_to_js(w0), steps=steps, targets=tg, law=lawd,
               rolloutSteps=rsteps, stateStride=stride, respawn=resp)
    r = run_node("gpu", job, tmp_path, timeout=240)

    # the same through the C ABI
    w = mk()
    g = G.create(w)
    G.upload(g)
    for s in range(steps):
        t = tg[s]
        ang = w.type == K.KX_ANGLE
        w.target[ang] = np.array([t[o % len(t)] for o in w.owner[ang]])
        G.upload_controls(g)
        G.step(g)
    st = G.download(g)
    G.destroy(g)
    h = r["host"]
    assert h["color"] == w.color.tolist() and h["a"] == w.a.tolist()
    for key in ("px", "py", "vx", "vy", "ang", "av"):
        assert np.array_equal(np.array(h[key]), getattr(w, key)), key
    assert np.array_equal(js_floats(h["lambda"]), w.lambda_)
    assert h["broken"] == w.broken.astype(int).tolist()
    assert h["stats"]["maxSpeed"] == st["maxSpeed"] and h["stats"]["nan"] == st["nan"]
    print(f"    node host loop identical over {steps} steps, dispatches {h['dispatches']}", flush=True)

    w2 = mk()
    g2 = G.create(w2)
    G.upload(g2)
    ev = G.rollout(g2, law, 0, rsteps, events=True)
    G.download(g2)
    rs = G.read_state(g2, stride)
    G.destroy(g2)
    ro = r["rollout"]
    for key in ("px", "py", "ang"):
        assert np.array_equal(np.array(ro[key]), getattr(w2, key)), key
    assert ro["events"] == ev.reshape(-1).tolist()
    assert ro["stateStride"] == rs.shape[1]
    assert np.array_equal(np.array(ro["state"], dtype=np.float32).reshape(rs.shape), rs)
    print(f"    node rollout identical over {rsteps} steps, {int((ev != 0).sum())} event words", flush=True)
    assert r["badThrows"] is True and "kxpbd error" in r["badMessage"]

    # respawn + substep through Node == through the C ABI
    w3 = mk()
    g3 = G.create(w3)
    G.upload(g3)
    G.rollout(g3, law, 0, resp["warm"])
    G.respawn(g3, resp["creatures"], np.array(resp["poses"]).reshape(-1, 3))
    G.substep(g3)
    G.download(g3)
    G.destroy(g3)
    for key in ("px", "py", "vx", "vy", "ang", "av"):
        assert np.array_equal(np.array(r["respawn"][key]), getattr(w3, key)), key
    assert r["respawn"]["broken"] == w3.broken.astype(int).tolist()
    # handle lifetime
    assert "busy" in r["busy"] and "not a live handle" in r["dead"] and r["drained"] is True


@needs_node
def test_brain_module_exports_the_net_and_es_surface():
    code = ("const b=require(process.argv[1]);"
            "console.log(JSON.stringify({k:Object.keys(b),w:b.weightCount(105),s:b.trainerStreamState(7)}))")
    r = subprocess.run([NODE, "-e", code, os.path.join(ROOT, "node", "brain.js")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    for name in ("inputDim", "weightCount", "createPopulation", "setGenomes", "getGenomes", "forwardInto",
                 "createES", "beginGeneration", "update", "esState", "trainerStreamState", "destroy"):
        assert name in d["k"], name
    assert d["w"] == 3644 and d["s"] == (7 * 2654435761) & 0xFFFFFFFF


@pytest.mark.gpu
@needs_node
def test_node_brain_matches_c_abi(tmp_path):
    """forwardInto over a population and two ES generations through node/brain.js
    give exactly what the same calls give through the C ABI from Python."""
    from kraken_3_gpu_amd import brain as B
    assert os.path.exists(ADDON), "node/kxpbd.node must be built"
    n, n_in, hidden = 6, B.input_dim(), B.HIDDEN
    nw = B.weight_count(n_in)
    rng = np.random.default_rng(42)
    gen = (rng.standard_normal((n, nw)) * 3 / np.sqrt(n_in)).astype(np.float32)
    xs = [rng.uniform(-1.5, 1.5, (n, n_in)).astype(np.float32) for _ in range(3)]
    theta = (rng.standard_normal(nw) * 0.1).astype(np.float32)
    fits = [rng.normal(0, 10, 4).tolist() for _ in range(2)]
    es_opts = dict(popSize=4, sigma=0.05, lr=0.02)
    job = dict(n=n, nIn=n_in, hidden=hidden, genomes=gen.reshape(-1).tolist(),
               inputs=[x.reshape(-1).tolist() for x in xs], theta=theta.tolist(), esOpts=es_opts, seed=9,
               fitness=fits)
    r = run_node("brain", job, tmp_path, timeout=120)
    assert r["nW"] == nw
    with B.Population(n, n_in) as P:
        P.set_genomes(gen)
        prev = np.zeros((n, B.OUT_DIM), np.float32)
        for x, st in zip(xs, r["steps"]):
            x = x.copy()
            x[:, n_in - B.OUT_DIM:] = prev
            out, ctl = P.forward_into(x, arms=6)
            assert np.array_equal(np.array(st["out"], np.float32).reshape(out.shape), out)
            assert np.array_equal(np.array(st["controls"], np.float32).reshape(ctl.shape), ctl)
            prev = out
    with B.ES(nw, theta=theta, **es_opts) as es, B.Population(4, n_in) as P:
        state = B.trainer_stream_state(9)
        for f, g in zip(fits, r["gens"]):
            state = es.begin_generation(P, state)
            assert g["state"] == state
            assert np.array_equal(np.array(g["genomes"], np.float32).reshape(4, nw), P.genomes())
            stats = es.update(np.array(f))
            s = es.state()
            for k in ("theta", "m", "v"):
                assert np.array_equal(np.array(g[k], np.float32), s[k]), k
            assert g["t"] == s["t"] and g["sigma"] == s["sigma"]
            for k in ("best", "mean", "sd", "sigma"):
                assert g["stats"][k] == stats[k], k
    assert "not a live policy" in r["deadThrows"]
    assert "does not match" in r["shapeThrows"] and "does not match" in r["popThrows"]
    print("    node brain identical: 3 forward steps, 2 ES generations", flush=True)


@needs_node
def test_brain_module_errors_without_a_gpu_call():
    """Shape checks and foreign handles throw with the library's message before
    any device work, as gpuworld.js throws."""
    code = ("const b=require(process.argv[1]);const out={};"
            "try{b.createPopulation(4,105,65);out.wide='no throw'}catch(e){out.wide=e.message}"
            "try{b.createES(3644,{popSize:9000});out.pop='no throw'}catch(e){out.pop=e.message}"
            "try{b.forwardInto({handle:5,n:1,nIn:105,nOut:20,hidden:24},new Float32Array(105),new Float32Array(20));"
            "out.dead='no throw'}catch(e){out.dead=e.message}"
            "console.log(JSON.stringify(out))")
    r = subprocess.run([NODE, "-e", code, os.path.join(ROOT, "node", "brain.js")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert "at most 64" in d["wide"] and "kxpbd error" in d["wide"]
    assert "8192" in d["pop"]
    assert "not a live policy" in d["dead"]
