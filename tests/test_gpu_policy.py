"""Policy network and ES on the device (kx_policy_*, kx_es_*) against the
reference's own vectors (tests/golden/net.json, es.json, written by
src/brain/net.js and src/brain/es.js) and against the oracle at population
sizes the fixtures do not reach.

Bars: ES genomes, theta, m and v bit-exact (every float64 operation is the
reference's, in its order). Forward outputs and activations: sums are the
reference's exactly; tanh is the device's float64 tanh, not V8's, so an
activation may differ by one float32 ulp where the two float64 results round
to different float32 neighbours — every value within 1 ulp, and at most 1 in
10^4 values off at all (observed: none)."""
import base64

import numpy as np
import pytest

from kraken_3_gpu_amd import _lib as K
from kraken_3_gpu_amd import brain as B
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def say(*a):
    print("   ", *a, flush=True)


def _f32(s):
    return np.frombuffer(base64.b64decode(s), np.float32)


def _ulps(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


def _within_one_ulp(dev, ref, what):
    u = _ulps(dev, ref)
    off = int((u > 0).sum())
    say(f"{what}: {u.size} values, {off} off by one ulp, max {int(u.max())} ulp")
    assert u.max() <= 1, what
    assert off <= max(1, u.size // 10000), what


def _tape_of(rec, arms):
    """decode's rows [tone, (shoulder, curl, adhesion) x arms, echo] as
    control-tape records."""
    t = np.zeros((rec.shape[0], K.KX_TAPE_WORDS), np.float32)
    t[:, :1 + 3 * arms] = rec[:, :1 + 3 * arms]
    return t


def test_forward_matches_reference_vectors(golden):
    d = golden("net.json")
    n_in, nw, pop = d["nIn"], d["nW"], d["pop"]
    gen = _f32(d["genomes"]).reshape(pop, nw)
    print(f"[case] policy forward, reference vectors: {pop} genomes x {nw} weights", flush=True)
    with B.Population(pop, n_in) as P:
        assert P.n_weights == nw
        P.set_genomes(gen)
        assert np.array_equal(P.genomes(), gen)
        for fr in d["frames"]:
            x = _f32(fr["inputs"]).reshape(pop, n_in)
            out, ctl, sc = P.forward_into(x, arms=6, want_scratch=True)
            _within_one_ulp(out, _f32(fr["outputs"]).reshape(pop, -1), "outputs")
            _within_one_ulp(sc, _f32(fr["scratch"]).reshape(pop, -1), "h1|h2")
            assert np.array_equal(ctl, _tape_of(po.decode(out, 6), 6))
            ref = np.array(fr["controls"])
            assert np.abs(ctl[:, :19] - ref[:, :19]).max() <= 1.2e-7


@pytest.mark.parametrize("n,scale,hidden", [(4096, 1.0, 24), (1001, 12.0, 24), (777, 3.0, 40)])
def test_forward_matches_oracle_population(n, scale, hidden):
    """A whole population against the oracle (random genomes of randomGenome's
    scale, some large enough to saturate), with the recurrent tail fed back for
    three steps; decode's records against the oracle's decode, at the
    reference's shape, an odd population, and a wider hidden layer."""
    n_in = B.input_dim()
    rng = np.random.default_rng(n)
    nw = B.weight_count(n_in, hidden)
    gen = (rng.standard_normal((n, nw)) * scale / np.sqrt(n_in)).astype(np.float32)
    print(f"[case] policy forward vs oracle: {n} creatures, scale {scale}, hidden {hidden}", flush=True)
    with B.Population(n, n_in, hidden=hidden) as P:
        P.set_genomes(gen)
        prev = np.zeros((n, B.OUT_DIM), np.float32)
        for step in range(3):
            x = (rng.uniform(-1.5, 1.5, (n, n_in))).astype(np.float32)
            x[:, n_in - B.OUT_DIM:] = prev
            out, ctl, sc = P.forward_into(x, arms=6, want_scratch=True)
            ref, ref_sc = po.forward_into(gen, x, n_in, hidden=hidden)
            _within_one_ulp(out, ref, f"step {step} outputs")
            _within_one_ulp(sc, ref_sc, f"step {step} h1|h2")
            assert np.array_equal(ctl, _tape_of(po.decode(out, 6), 6))
            prev = out
        say(f"kernel {P.last_ms():.4f} ms for {n} creatures")


def test_forward_rejects_bad_shapes():
    with pytest.raises(K.KxError, match="at most 64"):
        B.Population(4, 105, hidden=65)
    with pytest.raises(K.KxError, match="LDS"):
        B.Population(4, 1000, hidden=64, n_out=64)
    with B.Population(2, 105) as P:
        with pytest.raises(K.KxError, match="decode of 9 arms"):
            P.forward_into(np.zeros((2, 105), np.float32), arms=9)


def test_es_generations_match_reference_vectors(golden):
    e = golden("es.json")
    nw, pop = e["nW"], e["popSize"]
    o = e["opts"]
    print(f"[case] ES vs reference vectors: pop {pop}, {nw} weights", flush=True)
    es = B.ES(nw, theta=_f32(e["theta0"]), sigma=e["gens"][0]["sigma"], lr=o["lr"], popSize=pop,
              minSigma=o["minSigma"], sigmaDecay=o["sigmaDecay"], adamBeta1=o["adamBeta1"],
              adamBeta2=o["adamBeta2"], adamEps=o["adamEps"], clip=o["clip"])
    state = e["s0"]
    with es, B.Population(pop, B.input_dim()) as P:
        assert P.n_weights == nw
        for g in e["gens"]:
            state2 = es.begin_generation(P, state)
            assert state2 == (state + pop // 2 * nw * 4 * 0x6D2B79F5) & 0xFFFFFFFF
            state = state2
            assert np.array_equal(P.genomes(), _f32(g["genomes"]).reshape(pop, nw))
            st = es.update(np.array(g["fitness"]))
            s = es.state()
            assert np.array_equal(s["theta"], _f32(g["theta"]))
            assert np.array_equal(s["m"], _f32(g["m"])) and np.array_equal(s["v"], _f32(g["v"]))
            assert s["t"] == g["t"]
            for k in ("best", "mean", "sd", "sigma"):
                assert st[k] == g["stats"][k], k
        say("two generations: genomes, theta, m, v, statistics bit-exact")


@pytest.mark.parametrize("pop", [512, 1030])
def test_es_population_matches_oracle(pop):
    """512 and 1030 members (256 and 515 mirrored pairs: whole batches of the
    update's loads and a remainder) over the reference's network size, three
    generations, against the oracle; bestTheta follows the best fitness."""
    nw = B.weight_count(B.input_dim())
    rng = np.random.default_rng(5)
    theta = (rng.standard_normal(nw) * 0.1).astype(np.float32)
    opts = dict(B.ES_DEFAULTS)
    opts["popSize"] = pop
    print(f"[case] ES vs oracle: pop {pop}, {nw} weights", flush=True)
    th, m, v, t = theta.copy(), np.zeros(nw, np.float32), np.zeros(nw, np.float32), 0
    state = B.trainer_stream_state(11)
    best_ever, best_theta = -np.inf, None
    with B.ES(nw, theta=theta, popSize=pop) as es, B.Population(pop, B.input_dim()) as P:
        for gen in range(3):
            ref_g, eps, ref_state = po.genome_generation(th, opts["sigma"], pop, state)
            state = es.begin_generation(P, state)
            assert state == ref_state
            assert np.array_equal(P.genomes(), ref_g)
            fit = rng.normal(0, 10, pop) + gen
            st = es.update(fit)
            th, m, v, t, sigma, ref_st = po.es_update(th, m, v, t, eps, fit, opts)
            opts["sigma"] = sigma
            s = es.state()
            assert np.array_equal(s["theta"], th) and np.array_equal(s["m"], m) and np.array_equal(s["v"], v)
            for k in ("best", "mean", "sd", "sigma"):
                assert st[k] == ref_st[k], k
            if ref_st["best"] > best_ever:
                best_ever, best_theta = ref_st["best"], th.copy()
            assert s["best_ever"] == best_ever and np.array_equal(s["best_theta"], best_theta)
    say(f"three generations of {pop} bit-exact against the oracle")


def test_es_rejects_misuse():
    nw = B.weight_count(B.input_dim())
    with B.ES(nw, popSize=8) as es:
        with pytest.raises(K.KxError, match="before any generation"):
            es.update(np.zeros(8))
        with B.Population(6, B.input_dim()) as P:
            with pytest.raises(K.KxError, match="policy of 6 genomes"):
                es.begin_generation(P, 1)


def test_drive_matches_the_host_loop():
    """kx_policy_drive (readout, inputs, forward, decode into the tape, one
    tape step on the device) against the same step done from the host
    (read_state, inputs in numpy, forward_into with decode, upload_tape, a
    one-step KX_LAW_TAPE rollout): bodies, constraints and outputs identical
    over 12 steps, with senses, on twin worlds."""
    from kraken_3_gpu_amd import gpuworld as G
    from kraken_3_gpu_amd import world as W
    n_cr = 64
    n_in = B.input_dim()
    nw = B.weight_count(n_in)
    rng = np.random.default_rng(8)
    gen = (rng.standard_normal((n_cr, nw)) * 2 / np.sqrt(n_in)).astype(np.float32)
    senses = [rng.uniform(-1, 1, (n_cr, B.SENSE_DIM)).astype(np.float32) for _ in range(12)]
    law = W.default_law(77)
    tlaw = W.default_law(77)
    tlaw.flags |= K.KX_LAW_TAPE
    print(f"[case] policy drive vs host loop: {n_cr} creatures, 12 steps", flush=True)
    wa, wb = W.world_uniform(n_cr, seed=4), W.world_uniform(n_cr, seed=4)
    ga, gb = G.create(wa), G.create(wb)
    G.upload(ga)
    G.upload(gb)
    with B.Population(n_cr, n_in) as Pa, B.Population(n_cr, n_in) as Pb:
        Pa.set_genomes(gen)
        Pb.set_genomes(gen)
        prev = np.zeros((n_cr, B.OUT_DIM), np.float32)
        for s in range(12):
            Pa.drive(ga, law, s, senses[s])
            st = G.read_state(gb)
            x = np.concatenate([st[:, :42], senses[s], prev], axis=1)
            out, ctl = Pb.forward_into(x, arms=6)
            G.upload_tape(gb, ctl[None])
            G.rollout(gb, tlaw, s, 1)
            prev = out
        G.download(ga)
        G.download(gb)
        for key in ("px", "py", "ang", "vx", "vy", "av"):
            assert np.array_equal(getattr(wa, key), getattr(wb, key)), key
        assert np.array_equal(wa.lambda_, wb.lambda_) and np.array_equal(wa.broken, wb.broken)
        # the drive's outputs are the next step's recurrent inputs: one more
        # forward from the same inputs on both sides agrees
        x = np.concatenate([G.read_state(gb)[:, :42], senses[0], prev], axis=1)
        assert np.array_equal(Pb.forward_into(x), Pa.forward_into(x))
    moved = float(np.abs(wa.px - W.world_uniform(n_cr, seed=4).px).max())
    say(f"12 driven steps identical to the host loop; max body travel {moved:.2f} px")
    with pytest.raises(K.KxError, match="proprio"):
        with B.Population(n_cr, n_in) as Pc:
            Pc.drive(ga, law, 12, np.zeros((n_cr, 10), np.float32))   # 42 + 10 + 20 != 105
    with pytest.raises(K.KxError, match="creatures for a world"):
        with B.Population(8, n_in) as Pd:
            Pd.drive(ga, law, 12)
    G.destroy(ga)
    G.destroy(gb)


def test_drive_and_readout_reuse_one_state_buffer():
    """The policy step reads the state at its own stride (66 floats per
    creature) and a caller's readout at another (54): alternating them keeps
    one readout buffer per handle, grown once, so device memory stays flat."""
    import ctypes as C
    from kraken_3_gpu_amd import gpuworld as G
    from kraken_3_gpu_amd import world as W
    n_cr = 4096
    n_in = B.input_dim()
    nw = B.weight_count(n_in)
    rng = np.random.default_rng(3)
    gen = (rng.standard_normal((n_cr, nw)) * 2 / np.sqrt(n_in)).astype(np.float32)
    law = W.default_law(5)
    g = G.create(W.world_uniform(n_cr, seed=2))
    G.upload(g)

    def free_bytes():
        G.sync(g)
        f, t = C.c_int64(), C.c_int64()
        K.check(K.lib().kx_mem_info(C.byref(f), C.byref(t)))
        return f.value

    print(f"[case] drive / read_state alternation, {n_cr} creatures", flush=True)
    with B.Population(n_cr, n_in) as P:
        P.set_genomes(gen)
        P.drive(g, law, 0)
        G.read_state(g)
        before = free_bytes()
        for s in range(1, 21):
            P.drive(g, law, s)
            G.read_state(g)
        after = free_bytes()
    G.destroy(g)
    say(f"free device memory before {before / 2**20:.1f} MiB, after 20 alternations {after / 2**20:.1f} MiB")
    # a buffer re-allocated per call would hold 20 x ~1 MiB more
    assert before - after < 2 * 2**20
