"""Shared test plumbing: product worlds, their oracle twins, state transfer."""
import numpy as np

from oracle import pyoracle as po


def f32(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


def oracle_twin(w, f32_map=True):
    """Oracle world with the same creatures as product world `w` (built
    before gpuworld.create reorders it). The map uses the f32 rect values the
    device sees."""
    o = po.World(w.S, w.I, w.h)
    rects = w.rects.astype(np.float64) if f32_map else None
    o.set_map(w.width, w.height, rects)
    for (x, y, a, arms, segs, ring, curl) in w.spawns:
        o.spawn(x, y, a, arms, segs, ring, curl)
    o.color()
    return o


def load_state(o, w):
    """Copy the product world's (f32-valued) dynamic state into the oracle,
    whose constraint list order equals the handle order."""
    o.set_bodies(f32(w.px), f32(w.py), f32(w.ang), f32(w.vx), f32(w.vy), f32(w.av))
    o.set_cons_dynamic(broken=w.broken.astype(np.int32), lbx=f32(w.lbx), lby=f32(w.lby),
                       compliance=f32(w.compliance), target=f32(w.target), breakForce=f32(w.breakForce))
    cd = np.zeros((w.n_creatures, 8))
    cd[:, :] = w.cd
    o.set_cd(cd)


def decode_events(words):
    """[steps, creatures] u32 words -> sorted (step, creature, arm, kind)."""
    out = []
    steps, ncr = words.shape
    for s in range(steps):
        for c in np.nonzero(words[s])[0]:
            v = int(words[s, c])
            for bit in range(24):
                if v >> bit & 1:
                    out.append((s, int(c), bit // 3, bit % 3))
    return sorted(out)


def creature_ranges(w):
    """Body range [lo, hi) of every creature, from the recorded spawns."""
    out, lo = [], 0
    for (_, _, _, arms, segs, _, _) in w.spawns:
        n = 1 + arms * segs
        out.append((lo, lo + n))
        lo += n
    return out


def creature_of_bodies(w):
    cr = np.zeros(w.n, dtype=np.int64)
    for i, (lo, hi) in enumerate(creature_ranges(w)):
        cr[lo:hi] = i
    return cr


def oracle_subset(w, creatures):
    """Oracle world holding only `creatures` of product world `w`, plus the
    index maps (bodies, constraints in handle order) into w's arrays."""
    o = po.World(w.S, w.I, w.h)
    o.set_map(w.width, w.height, w.rects.astype(np.float64))
    rng = creature_ranges(w)
    bidx = []
    for c in creatures:
        x, y, a, arms, segs, ring, curl = w.spawns[c]
        o.spawn(x, y, a, arms, segs, ring, curl)
        bidx.extend(range(*rng[c]))
    o.color()
    crb = creature_of_bodies(w)
    keep = np.isin(crb[w.a], np.asarray(creatures))
    cidx = np.nonzero(keep)[0]
    return o, np.asarray(bidx), cidx


def load_state_subset(o, w, creatures, bidx, cidx):
    o.set_bodies(*(f32(getattr(w, k)[bidx]) for k in ("px", "py", "ang", "vx", "vy", "av")))
    o.set_cons_dynamic(broken=w.broken[cidx].astype(np.int32), lbx=f32(w.lbx[cidx]), lby=f32(w.lby[cidx]),
                       compliance=f32(w.compliance[cidx]), target=f32(w.target[cidx]),
                       breakForce=f32(w.breakForce[cidx]))
    o.set_cd(np.ascontiguousarray(w.cd[np.asarray(creatures)]))
