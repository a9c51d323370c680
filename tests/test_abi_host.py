"""Host-side checks of libkxpbd (no GPU needed): ABI exports, the product
world builder against the oracle's, colouring, point grouping, the pure
gpuworld helpers against the reference's own outputs, and error behaviour."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from kraken_3_gpu_amd import _lib as K
from kraken_3_gpu_amd import gpuworld as G
from kraken_3_gpu_amd import world as W
from oracle import pyoracle as po

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "kxpbd.h")).read()
    names = set(re.findall(r"\b(kx_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) > 25
    L = K.lib()
    for n in sorted(names):
        assert hasattr(L, n), n
    assert L.kx_abi_version() == 5


def _oracle_like(w):
    o = po.World(w.S, w.I, w.h)
    for (x, y, a, arms, segs, ring, curl) in w.spawns:
        o.spawn(x, y, a, arms, segs, ring, curl)
    return o


@pytest.mark.parametrize("maker", ["cpu64", "uneven"])
def test_builder_matches_oracle_builder(maker):
    w = W.world_cpu64(8) if maker == "cpu64" else W.world_uneven(12, sort_by_size=False)
    o = _oracle_like(w)
    b = o.bodies()
    assert np.array_equal(w.px, b["px"]) and np.array_equal(w.py, b["py"]) and np.array_equal(w.ang, b["ang"])
    assert np.array_equal(w.im, b["im"].astype(np.float32)) and np.array_equal(w.ii, b["ii"].astype(np.float32))
    c = o.cons()
    for key, mine in (("type", w.type), ("a", w.a), ("b", w.b), ("owner", w.owner)):
        assert np.array_equal(mine, c[key]), key
    for key, mine in (("lax", w.lax), ("lay", w.lay), ("rest", w.rest), ("minA", w.minA), ("maxA", w.maxA)):
        assert np.array_equal(mine, c[key].astype(np.float32).astype(np.float64)), key
    for key, mine in (("lbx", w.lbx), ("lby", w.lby), ("compliance", w.compliance), ("breakForce", w.breakForce)):
        assert np.array_equal(mine, c[key]), key
    assert np.array_equal(w.broken, c["broken"].astype(np.uint8))
    p = o.points()
    assert np.array_equal(w.p_body, p["body"])
    assert np.array_equal(w.p_r, p["r"].astype(np.float32))


def test_builder_map_matches_reference(golden):
    for m in golden("maps.json")["maps"]:
        b = W.Builder()
        o = m["opts"]
        b.map(m["width"], m["height"], o["seed"], o.get("corridorCols", 4), o.get("corridorRows", 3))
        w = b.world()
        assert np.array_equal(w.rects, np.array(m["rects"], dtype=np.float32).reshape(-1, 4))


def _color(w):
    col = np.zeros(w.nc, dtype=np.int32)
    order = np.zeros(w.nc, dtype=np.int32)
    K.check(K.lib().kx_color_constraints(w.n, w.nc, K.ptr(w.a, C.c_int32), K.ptr(w.b, C.c_int32), None,
                                         K.ptr(col, C.c_int32), K.ptr(order, C.c_int32)))
    return col, order


def test_coloring_matches_oracle_and_reference(golden):
    w = W.world_uneven(10, sort_by_size=False)
    col, order = _color(w)
    o = _oracle_like(w)
    o.color()
    c = o.cons()
    assert np.array_equal(order, c["id"])
    assert np.array_equal(col[order], c["color"])
    for case in golden("kraken.json")["cases"]:
        x, y, a, nk = case["spawn"]
        b = W.Builder()
        b.map(1200, 800, 1)
        for i in range(nk):
            b.spawn(x + 250 * i, y - 40 * i, a + 0.3 * i)
        wk = b.world()
        col, order = _color(wk)
        assert order.tolist() == case["sorted"]
        assert col[order].tolist() == case["colorOf"]
        cnt = np.bincount(col, minlength=16)
        assert cnt.tolist() == case["colorCnt"]


def test_group_points_matches_reference(golden):
    g = golden("gpuworld.json")["group"]
    nB = golden("gpuworld.json")["nB"]
    order, rng = G.group_coll_points(g["bodies"], nB)
    assert order.tolist() == g["order"] and rng.tolist() == g["range"]
    order, rng = G.group_coll_points(g["shuffledBodies"], nB)
    assert order.tolist() == g["shuffledOrder"] and rng.tolist() == g["shuffledRange"]
    with pytest.raises(K.KxError):
        G.group_coll_points([0, 5], 3)


def _sim_world(seed, creatures):
    """createSim({seed, creatures}) layout (core.js:37-74) in colour order."""
    b = W.Builder()
    b.map(1440, 810, seed)
    st = [(seed ^ 0x9E3779B9) & 0xFFFFFFFF]
    for _ in range(creatures):
        xy = b.find_spawn(W.ARM_REACH + 20, st)
        b.spawn(xy[0], xy[1], W._rng_next(st) * math.pi * 2)
    w = b.world(dict(substeps=2, iterations=4, h=1 / 120))
    col, order = _color(w)
    w.color = col
    w.reorder_constraints(order)
    return w


def test_pack_controls_reference_layout(golden):
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
    cd = np.zeros(w.nc * G.GPU_STRIDES["constraint"], dtype=np.float32)
    G.pack_controls(w, cd)
    assert cd.view(np.uint32).tolist() == g["packed"]


def test_mirror_into_reference_fields(golden):
    g = golden("gpuworld.json")
    w = _sim_world(33, 2)
    bd = np.array(g["bodyIn"], dtype=np.uint32).view(np.float32)
    cdat = np.array(g["consIn"], dtype=np.uint32).view(np.float32)
    stats = G.mirror_into(w, bd, cdat)
    m = g["mirrored"]
    px = [float("nan") if v == "NaN" else v for v in m["px"]]
    py = [float("nan") if v == "NaN" else v for v in m["py"]]
    np.testing.assert_array_equal(w.px, px)
    np.testing.assert_array_equal(w.py, py)
    assert w.lambda_.tolist() == m["lambda"]
    assert w.broken.tolist() == m["broken"]
    grip = w.type == K.KX_GRIP
    assert w.force[grip].tolist() == np.array(m["force"])[grip].tolist()
    assert stats["maxSpeed"] == m["stats"]["maxSpeed"] and stats["nan"] == m["stats"]["nan"]
    assert stats["maxAngSpeed"] == m["stats"]["maxAngSpeed"]


def test_v8_hypot_is_bit_exact():
    assert G.v8_hypot(3.0, 4.0) == 5.0
    rng = np.random.default_rng(5)
    x, y = rng.normal(size=2000) * 500, rng.normal(size=2000) * 500
    mine = G.v8_hypot(x, y)
    ref = np.array([po.lib().or_js_hypot(a, b) for a, b in zip(x, y)])
    assert np.array_equal(mine, ref)


def _desc_for(w, keep, **over):
    d, c = w.desc(keep)
    for k, v in over.items():
        setattr(d, k, v)
    return d, c


@pytest.mark.parametrize("case", ["iterations", "h", "rects", "index"])
def test_create_sanity_guards_raise(case):
    """gpuworld.js:224-229 and 269-271 throw before touching the device."""
    w = W.world_cpu64(2)
    keep = []
    d, c = w.desc(keep)
    prm = K.Params(4, 8, 1 / 240, 3.0, 3.0, 360.0, 18.0, K.KX_PATH_PERSISTENT, 0, 0)
    if case == "iterations":
        prm.iterations = 0
    elif case == "h":
        prm.h = 1.5
    elif case == "rects":
        d.n_rects = 0
    else:
        bad = w.b.copy()
        bad[3] = w.n + 7
        keep.append(bad)
        d.c_b = K.ptr(bad, C.c_int32)
    h = C.c_void_p()
    rc = K.lib().kx_create(C.byref(d), C.byref(c), C.byref(prm), C.byref(h))
    assert rc == K.KX_EINVAL and not h.value
    assert K.lib().kx_last_error()


def test_policy_shapes_and_argument_checks_on_the_host():
    """weightCount (src/brain/net.js:30-32) and the argument checks that run
    before any device call."""
    import ctypes as C
    L = K.lib()
    assert L.kx_policy_weights(105, 24, 20) == 3644 == po.weight_count(105)
    assert L.kx_policy_weights(0, 24, 20) == 0
    h = C.c_void_p()
    assert L.kx_policy_create(0, 105, 24, 20, C.byref(h)) == K.KX_EINVAL
    assert L.kx_policy_create(4, 105, 65, 20, C.byref(h)) == K.KX_EINVAL
    assert "at most 64" in L.kx_last_error().decode()
    assert L.kx_policy_create(4, 1000, 64, 64, C.byref(h)) == K.KX_EINVAL
    assert "LDS" in L.kx_last_error().decode()
    eo = K.EsOpts(sigma=0.06, lr=0.03, min_sigma=0.004, sigma_decay=0.999, beta1=0.9, beta2=0.999,
                  adam_eps=1e-8, clip=2.5, pop_size=0)
    assert L.kx_es_create(3644, C.byref(eo), None, C.byref(h)) == K.KX_EINVAL
    assert L.kx_es_create(0, C.byref(eo), None, C.byref(h)) == K.KX_EINVAL
    eo.pop_size = 8194
    assert L.kx_es_create(3644, C.byref(eo), None, C.byref(h)) == K.KX_EINVAL
    assert "8192" in L.kx_last_error().decode()
