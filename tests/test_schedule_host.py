"""The persistent path's round schedule (host planning, no GPU): structure
checks and bit-exact equivalence with colour-order Gauss-Seidel on the
oracle (src/physics/xpbd.js:290-296 iterates the colour-sorted list)."""
import ctypes as C

import numpy as np
import pytest

from kraken_3_gpu_amd import _lib as K
from kraken_3_gpu_amd import world as W
from oracle import pyoracle as po


KX_PLAN_SINGLE_LANE = 0x100


def plan(nb, a, b, t, G, single=False):
    a, b, t = (np.ascontiguousarray(x, dtype=np.int32) for x in (a, b, t))
    sched = np.zeros(4096 * G, dtype=np.int32)
    rtype = np.zeros(4096, dtype=np.int32)
    g = G | (KX_PLAN_SINGLE_LANE if single else 0)
    R = K.check(K.lib().kx_plan_rounds(nb, len(a), K.ptr(a, C.c_int32), K.ptr(b, C.c_int32), K.ptr(t, C.c_int32), g,
                                       K.ptr(sched, C.c_int32), K.ptr(rtype, C.c_int32), 4096))
    return sched[: R * G].reshape(R, G), rtype[:R]


def creature_cons(arms=6, segs=7, ring=8):
    o = po.World()
    o.spawn(600, 400, 0.3, arms, segs, ring)
    o.color()
    c = o.cons()
    return o.n, c["a"], np.where(c["b"] >= 0, c["b"], -1), c["type"]


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("G", [8, 16, 32, 64])
@pytest.mark.parametrize("morph", [(6, 7, 8), (4, 5, 6), (8, 9, 10), (5, 8, 7)])
def test_rounds_are_valid_topological_antichains(G, morph, single):
    nb, a, b, t = creature_cons(*morph)
    sched, rtype = plan(nb, a, b, t, G, single)
    # a pin round holds a lane pair per pin, or one lane per pin when single
    for r, row in enumerate(sched):
        if not rtype[r]:
            assert (row >= 0).sum() <= (G if single else G // 2)
    flat = sched[sched >= 0]
    assert sorted(flat.tolist()) == list(range(len(a)))
    where = np.empty(len(a), dtype=np.int64)
    for r, row in enumerate(sched):
        ks = row[row >= 0]
        assert len(ks) >= 1
        assert all((t[k] == 1) == bool(rtype[r]) for k in ks)
        bodies = np.concatenate([a[ks], b[ks][b[ks] >= 0]])
        assert len(bodies) == len(set(bodies.tolist()))
        where[ks] = r
    for k in range(len(a)):
        for j in range(k):
            share = {a[k], b[k]} & {a[j], b[j]} - {-1}
            if share:
                assert where[j] < where[k]
    # critical path of the 6x7 creature: the mantle's 12 constraints chain
    if morph == (6, 7, 8):
        assert len(sched) >= 12


def test_round_counts_shrink_with_lanes():
    nb, a, b, t = creature_cons()
    R = {G: len(plan(nb, a, b, t, G)[0]) for G in (8, 16, 32, 64)}
    assert R[64] <= R[32] <= R[16] <= R[8]
    assert R[16] <= 20 and R[8] <= 24


@pytest.mark.parametrize("G", [8, 16])
def test_schedule_order_is_bit_identical_to_colour_order(G):
    """Stepping the oracle with its constraint list in round order gives
    exactly the colour-order result (commuting body-disjoint constraints)."""
    law_seed = 4
    w = W.world_cpu64(3, seed=2)
    ref = po.World(w.S, w.I, w.h)
    alt = po.World(w.S, w.I, w.h)
    for o in (ref, alt):
        o.set_map(w.width, w.height, w.rects.astype(np.float64))
        for sp in w.spawns:
            o.spawn(*sp)
        o.color()
    c = ref.cons()
    nb = ref.n
    sched, _ = plan(nb, c["a"], np.where(c["b"] >= 0, c["b"], -1), c["type"], G)
    order = sched[sched >= 0]
    alt.permute_cons(order)
    for s in range(30):
        ref.apply_law(law_seed, s)
        alt.apply_law(law_seed, s)
        ref.step()
        alt.step()
    br, ba = ref.bodies(), alt.bodies()
    for key in ("px", "py", "ang", "vx", "vy", "av"):
        assert np.array_equal(br[key], ba[key]), key
    assert sorted(ref.events()) == sorted(alt.events())


@pytest.mark.parametrize("morph", [(4, 5, 6), (5, 6, 7), (6, 7, 8)])
def test_single_lane_pins_alternate_at_8_lanes(morph):
    """With a pin on each lane an 8-lane round holds eight pins, so these
    shapes get 2 x arms rounds alternating pins and motors (what the fused
    rounds need), where lane pairs need extra pin rounds."""
    nb, a, b, t = creature_cons(*morph)
    _, rt_pair = plan(nb, a, b, t, 8)
    _, rt_single = plan(nb, a, b, t, 8, single=True)
    arms = morph[0]
    assert len(rt_single) == 2 * arms
    assert all(rt_single[r] == r % 2 for r in range(len(rt_single)))
    assert len(rt_pair) > 2 * arms
