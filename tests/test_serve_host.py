"""Host side of the serving loop (no GPU): the policy's control tape against
a numpy restatement, and the request stream generator."""
import math

import numpy as np

from kraken_3_gpu_amd import _lib as K
from kraken_3_gpu_amd import serve as S
from kraken_3_gpu_amd import world as W

NI, NH, NO = 64, 32, 48
f32 = np.float32


def _gauss(rnd):
    """gauss (src/brain/net.js:56-62)."""
    u = v = 0.0
    while u == 0.0:
        u = rnd()
    while v == 0.0:
        v = rnd()
    return math.sqrt(-2.0 * math.log(u)) * math.cos(2.0 * math.pi * v)


def _theta(seed):
    r = W.mulberry32(seed)
    th = np.zeros(NI * NH + NH + NH * NO + NO, np.float32)
    s1, s2 = 1.0 / math.sqrt(NI), 1.0 / math.sqrt(NH)
    for i in range(NI * NH):
        th[i] = _gauss(r) * s1
    for k in range(NH * NO):
        th[NI * NH + NH + k] = _gauss(r) * s2
    return th


def _eps(noise_seed, p):
    r = W.mulberry32((noise_seed ^ (p * 0x9E3779B9)) & 0xFFFFFFFF)
    r()
    seed = int(r() * 4294967296.0)
    r = W.mulberry32(seed)
    return np.array([_gauss(r) for _ in range(NI * NH + NH + NH * NO + NO)], np.float32)


def _tanh32(x):
    """The tape's binary32 tanh, restated: odd Taylor series below 1/16, else
    1 - 2/(exp(2|x|)+1) with a ln2 range reduction and degree-6 Taylor exp."""
    x = np.asarray(x, np.float32)
    ax = np.abs(x)
    x2 = x * x
    small = x * (f32(1) + x2 * (f32(-1.0 / 3.0) + x2 * (f32(2.0 / 15.0) + x2 * f32(-17.0 / 315.0))))
    y = np.minimum(f32(2) * ax, f32(18))
    k = np.rint(y * f32(1.44269504088896340736))
    r = (y - k * f32(0.693145751953125)) - k * f32(1.42860682030941723212e-6)
    p = f32(1) + r * (f32(1) + r * (f32(0.5) + r * (f32(1.0 / 6.0) + r * (f32(1.0 / 24.0) + r * (
        f32(1.0 / 120.0) + r * f32(1.0 / 720.0))))))
    e = p * np.ldexp(f32(1), k.astype(np.int32)).astype(np.float32)
    big = f32(1) - f32(2) / (e + f32(1))
    big = np.where(ax >= f32(9), f32(1), big).astype(np.float32)
    big = np.copysign(big, x)
    return np.where(ax < f32(0.0625), small, big).astype(np.float32)


def _tape(d, step0, n, step_s):
    th = _theta(d["policy_seed"])
    W1 = th[:NI * NH].reshape(NH, NI)
    b1 = th[NI * NH:NI * NH + NH]
    W2 = th[NI * NH + NH:NI * NH + NH + NH * NO].reshape(NO, NH)
    b2 = th[NI * NH + NH + NH * NO:]
    P, sig = d["pairs"], d["sigma"]
    out_t = np.zeros((n, 2 * P, K.KX_TAPE_WORDS), np.float32)
    eps = [_eps(d["noise_seed"], p) for p in range(P)]
    for r in range(n):
        t = (step0 + r) * step_s
        x = np.zeros(NI, np.float32)
        for k in range(NI // 2):
            w = 2.0 * math.pi * 0.125 * (k + 1) * t
            x[2 * k], x[2 * k + 1] = math.sin(w), math.cos(w)
        A = b1.copy()
        for i in range(NI):
            A = (A + x[i] * W1[:, i]).astype(np.float32)
        for p in range(P):
            e = eps[p]
            E1 = e[:NI * NH].reshape(NH, NI)
            EB1 = e[NI * NH:NI * NH + NH]
            E2 = e[NI * NH + NH:NI * NH + NH + NH * NO].reshape(NO, NH)
            EB2 = e[NI * NH + NH + NH * NO:]
            D = (sig * EB1.astype(np.float64)).astype(np.float32)
            sE1 = (sig * E1.astype(np.float64)).astype(np.float32)
            for i in range(NI):
                D = (D + x[i] * sE1[:, i]).astype(np.float32)
            for sgn, hpre in ((0, A + D), (1, A - D)):
                s = -sig if sgn else sig
                h = _tanh32(hpre.astype(np.float32))
                W2c = (W2.astype(np.float64) + E2.astype(np.float64) * s).astype(np.float32)
                o = (b2.astype(np.float64) + EB2.astype(np.float64) * s).astype(np.float32)
                for j in range(NH):
                    o = (o + h[j] * W2c[:, j]).astype(np.float32)
                o = _tanh32(o)
                rec = out_t[r, 2 * p + sgn]
                rec[0] = f32(0.5) + f32(0.5) * o[0]
                for a in range(d["arms"]):
                    rec[1 + 3 * a], rec[2 + 3 * a] = o[1 + 3 * a], o[2 + 3 * a]
                    rec[3 + 3 * a] = o[3 + 3 * a] if o[3 + 3 * a] > 0 else 0.0
    return out_t


def _req(**kw):
    d = dict(arms=5, segs=6, ring=7, pairs=3, steps=40, x=300.0, y=200.0, angle=0.5, policy_seed=11,
             noise_seed=77, sigma=0.06, tag=1)
    d.update(kw)
    return d


def test_policy_tape_matches_restatement():
    """kx_policy_tape (the records the server feeds the device) equals the
    numpy restatement bit for bit: seeded theta and eps (mulberry32 +
    Box-Muller), mirrored pairs, phase features, two tanh layers, decode."""
    step_s = 4 / 240
    for d, step0, n in ((_req(), 0, 3), (_req(arms=8, pairs=2, policy_seed=5, noise_seed=9), 117, 2)):
        got = S.policy_tape(d, step0, n, step_s)
        want = _tape(d, step0, n, step_s)
        assert got.shape == want.shape
        assert np.array_equal(got, want)


def test_policy_tape_properties():
    """decode ranges (tone in [0,1], adhesion >= 0, unused arms zero), sigma=0
    makes the two members of a pair identical, and the tape of steps a..b is
    the slice of a longer tape (rows depend only on the step)."""
    d = _req(pairs=4)
    t = S.policy_tape(d, 0, 30)
    assert t.shape == (30, 8, K.KX_TAPE_WORDS)
    assert (t[..., 0] >= 0).all() and (t[..., 0] <= 1).all()
    assert (t[..., 3:1 + 3 * 5:3] >= 0).all()
    assert (t[..., 1 + 3 * 5:] == 0).all()
    assert not np.array_equal(t[:, 0], t[:, 1])
    assert np.array_equal(S.policy_tape(d, 10, 5), t[10:15])
    z = S.policy_tape(dict(d, sigma=0.0), 0, 6)
    assert np.array_equal(z[:, 0::2], z[:, 1::2])
    assert np.abs(t[1:] - t[:-1]).max() > 0       # the features move the outputs


def test_tanh32_accuracy():
    x = np.linspace(-12, 12, 200001, dtype=np.float32)
    err = np.abs(_tanh32(x).astype(np.float64) - np.tanh(x.astype(np.float64)))
    assert err.max() < 3e-7


def test_es_stream_draws():
    """Seeded stream: sizes and steps in range, morphologies from the uneven
    range, arrivals increasing with the stated mean gap."""
    t, reqs = S.es_stream(400, rate=20.0, seed=3)
    assert len(reqs) == 400 and t[0] == 0.0 and (np.diff(t) >= 0).all()
    assert abs(np.diff(t).mean() - 1 / 20.0) < 0.01
    assert all(16 <= r["pairs"] <= 512 and 300 <= r["steps"] <= 1500 for r in reqs)
    assert all(4 <= r["arms"] <= 8 and 5 <= r["segs"] <= 9 and 6 <= r["ring"] <= 10 for r in reqs)
    t2, reqs2 = S.es_stream(400, rate=20.0, seed=3)
    assert np.array_equal(t, t2) and reqs == reqs2
