"""kx_rollout_noalloc: kx_rollout whose per-step event words are staged in
the library's static device array instead of its grow-on-demand events buffer
(for callers that cannot allocate). Checked against kx_rollout on twin
handles: events and final state identical, on two calls in a row (the second
on a smaller world and fewer steps, so stale words of the first would show),
free device memory unchanged across a call, and sizes beyond the array
refused."""
import ctypes as C
import sys

import numpy as np
import pytest

from kraken_3_gpu_amd import _lib as K
from kraken_3_gpu_amd import gpuworld as G
from kraken_3_gpu_amd import world as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc():
    return K.lib().kx_rollout_noalloc


def _free_bytes():
    f, t = C.c_int64(), C.c_int64()
    K.check(K.lib().kx_mem_info(C.byref(f), C.byref(t)))
    return f.value


@pytest.mark.parametrize("path", ["persistent", "plain"])
def test_noalloc_rollout_equals_rollout(path):
    p = {"persistent": K.KX_PATH_PERSISTENT, "plain": K.KX_PATH_PLAIN}[path]
    law = W.default_law(2024)
    fn = _noalloc()
    for call, (n_cr, step0, n) in enumerate(((32, 0, 40), (8, 0, 12))):
        wa, wb = W.world_cpu64(n_cr, seed=1), W.world_cpu64(n_cr, seed=1)
        ga, gb = G.create(wa, path=p), G.create(wb, path=p)
        G.upload(ga)
        G.upload(gb)
        say(f"[case] no-alloc rollout path={path} call={call + 1} creatures={n_cr} steps={n}")
        ea = G.rollout(ga, law, step0, n, events=True)
        eb = np.full((n, gb.nCr), 0xDEADBEEF, np.uint32)
        free0 = _free_bytes()
        K.check(fn(gb.p, C.byref(law), step0, n, K.ptr(eb, C.c_uint32)))
        free1 = _free_bytes()
        G.download(ga)
        G.download(gb)
        say(f"    events {int((ea != 0).sum())} words set; free device memory {free0} -> {free1}")
        assert np.array_equal(ea, eb) and (ea != 0).any()
        for k in ("px", "py", "ang", "vx", "vy", "av"):
            assert np.array_equal(getattr(wa, k), getattr(wb, k)), k
        assert np.array_equal(wa.broken, wb.broken)
        if call:
            assert free1 == free0
        G.destroy(ga)
        G.destroy(gb)


def test_noalloc_rollout_refuses_sizes_beyond_its_scratch():
    w = W.world_cpu64(8, seed=1)
    g = G.create(w)
    G.upload(g)
    law = W.default_law(1)
    ev = np.zeros(8, np.uint32)
    with pytest.raises(K.KxError, match="exceeds"):
        K.check(_noalloc()(g.p, C.byref(law), 0, (1 << 22) // 8 + 1, K.ptr(ev, C.c_uint32)))
    G.destroy(g)
