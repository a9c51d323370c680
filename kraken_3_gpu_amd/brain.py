"""Policy network and evolution strategy on the device, over the C ABI
(kx_policy_*, kx_es_*): the batched face of src/brain/net.js and
src/brain/es.js as src/sim/gputrain.js:55-86 and src/sim/train.js:154-158
drive them — one genome per creature in one flat array, one forward pass for
the whole population per step, one ES generation and update per episode.

Every call runs the HIP library; there is no host fallback."""
import ctypes as C

import numpy as np

from . import _lib as K

HIDDEN = 24                      # src/brain/net.js:20
OUT_DIM = 1 + 6 * 3 + 1          # src/brain/net.js:23
PROPRIO_DIM = 42                 # src/sim/core.js:123
SENSE_DIM = 43                   # src/sim/core.js:126 (echo, chemical, light, hearing, 6 more)

# es.js:23-38 defaults
ES_DEFAULTS = dict(sigma=0.06, lr=0.030, popSize=64, minSigma=0.004, sigmaDecay=0.999, adamBeta1=0.9,
                   adamBeta2=0.999, adamEps=1e-8, clip=2.5)


def input_dim(proprio_dim=PROPRIO_DIM, sense_dim=SENSE_DIM):
    """src/brain/net.js:25-27."""
    return proprio_dim + sense_dim + OUT_DIM


def weight_count(n_in, hidden=HIDDEN, n_out=OUT_DIM):
    """src/brain/net.js:30-32 (kx_policy_weights)."""
    return K.lib().kx_policy_weights(n_in, hidden, n_out)


def trainer_stream_state(seed):
    """State of the trainer's stream before its first call
    (src/sim/train.js:65: (seed * 2654435761) >>> 0)."""
    return (int(seed) * 2654435761) & 0xFFFFFFFF


class Population:
    """n creatures' genomes resident on the device and the batched forwardInto
    (src/brain/net.js:71-98) over them."""

    def __init__(self, n, n_in=None, hidden=HIDDEN, n_out=OUT_DIM):
        self.n = int(n)
        self.n_in = int(n_in if n_in is not None else input_dim())
        self.hidden, self.n_out = int(hidden), int(n_out)
        self.n_weights = weight_count(self.n_in, self.hidden, self.n_out)
        h = C.c_void_p()
        K.check(K.lib().kx_policy_create(self.n, self.n_in, self.hidden, self.n_out, C.byref(h)))
        self.p = h

    def close(self):
        if self.p:
            K.lib().kx_policy_destroy(self.p)
            self.p = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_genomes(self, genomes):
        g = np.ascontiguousarray(genomes, np.float32).reshape(self.n, self.n_weights)
        K.check(K.lib().kx_policy_set_genomes(self.p, K.ptr(g, C.c_float)))

    def genomes(self):
        g = np.zeros((self.n, self.n_weights), np.float32)
        K.check(K.lib().kx_policy_get_genomes(self.p, K.ptr(g, C.c_float)))
        return g

    def forward_into(self, x, arms=None, want_scratch=False):
        """x float32 [n, n_in] (previous outputs in the last n_out columns, as
        gputrain.js:82 places them) -> out float32 [n, n_out]; with `arms`,
        also decode's control-tape records float32 [n, KX_TAPE_WORDS]
        (src/brain/net.js:114-124); with want_scratch, h1|h2 [n, 2*hidden]."""
        x = np.ascontiguousarray(x, np.float32).reshape(self.n, self.n_in)
        out = np.zeros((self.n, self.n_out), np.float32)
        sc = np.zeros((self.n, 2 * self.hidden), np.float32) if want_scratch else None
        ctl = np.zeros((self.n, K.KX_TAPE_WORDS), np.float32) if arms is not None else None
        K.check(K.lib().kx_policy_forward(self.p, K.ptr(x, C.c_float), K.ptr(out, C.c_float),
                                          K.ptr(sc, C.c_float), K.ptr(ctl, C.c_float),
                                          int(arms) if arms is not None else 0))
        res = [out]
        if arms is not None:
            res.append(ctl)
        if want_scratch:
            res.append(sc)
        return res[0] if len(res) == 1 else tuple(res)

    def drive(self, g, law, step, senses=None):
        """One step of src/sim/gputrain.js:77-100 on the device for the world
        handle g (gpuworld.create): readProprio, inputs [proprio | senses |
        previous outputs], forwardInto, decode into the control tape, one
        rollout step with applyControl (kx_policy_drive). senses float32
        [n, sense_dim] or None (zeros, sense_dim = n_in - 42 - n_out at six
        arms)."""
        if senses is None:
            sd = self.n_in - self.n_out - 42
            sp = None
        else:
            senses = np.ascontiguousarray(senses, np.float32).reshape(self.n, -1)
            sd, sp = senses.shape[1], K.ptr(senses, C.c_float)
        K.check(K.lib().kx_policy_drive(self.p, g.p, sp, int(sd), C.byref(law), int(step)))

    def last_ms(self):
        v = C.c_float()
        K.check(K.lib().kx_policy_last_ms(self.p, C.byref(v)))
        return v.value

    def forward_repeat(self, reps):
        """Mean device time (ms) of `reps` forward kernels over the resident
        inputs and genomes (measurement only)."""
        v = C.c_float()
        K.check(K.lib().kx_policy_forward_repeat(self.p, int(reps), C.byref(v)))
        return v.value


class ES:
    """createES / genomeAt / update (src/brain/es.js:22-149) with theta, Adam
    moments and the eps bank on the device."""

    def __init__(self, n_weights, theta=None, **opts):
        o = dict(ES_DEFAULTS)
        o.update(opts)
        self.opts = o
        eo = K.EsOpts(sigma=o["sigma"], lr=o["lr"], min_sigma=o["minSigma"], sigma_decay=o["sigmaDecay"],
                      beta1=o["adamBeta1"], beta2=o["adamBeta2"], adam_eps=o["adamEps"], clip=o["clip"],
                      pop_size=int(o["popSize"]))
        self.n_weights = int(n_weights)
        self.pop_size = int(o["popSize"]) + (int(o["popSize"]) & 1)
        th = None if theta is None else np.ascontiguousarray(theta, np.float32).reshape(self.n_weights)
        h = C.c_void_p()
        K.check(K.lib().kx_es_create(self.n_weights, C.byref(eo), K.ptr(th, C.c_float), C.byref(h)))
        self.e = h

    def close(self):
        if self.e:
            K.lib().kx_es_destroy(self.e)
            self.e = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def begin_generation(self, pop, rnd_state):
        """genomeAt for i = 0 .. popSize-1 into `pop`'s resident genomes
        (src/sim/train.js:154-158) from the trainer's stream state; returns
        the state after the draws."""
        st = C.c_uint32(int(rnd_state) & 0xFFFFFFFF)
        K.check(K.lib().kx_es_generation(self.e, pop.p, C.byref(st)))
        return st.value

    def update(self, fitness):
        """src/brain/es.js:92-149; returns {best, mean, sd, sigma}."""
        f = np.ascontiguousarray(fitness, np.float64).reshape(self.pop_size)
        st = np.zeros(4, np.float64)
        K.check(K.lib().kx_es_update(self.e, K.ptr(f, C.c_double), K.ptr(st, C.c_double)))
        return dict(best=st[0], mean=st[1], sd=st[2], sigma=st[3])

    def state(self):
        """theta, m, v, bestTheta (float32 [nW]) and sigma, t, bestEver."""
        a = [np.zeros(self.n_weights, np.float32) for _ in range(4)]
        sc = np.zeros(3, np.float64)
        K.check(K.lib().kx_es_get(self.e, *[K.ptr(x, C.c_float) for x in a], K.ptr(sc, C.c_double)))
        return dict(theta=a[0], m=a[1], v=a[2], best_theta=a[3], sigma=sc[0], t=int(sc[1]), best_ever=sc[2])
