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


def _js_num(x):
    x = float(x)
    if x != x:
        return "NaN"
    if x == float("inf"):
        return "Infinity"
    if x == float("-inf"):
        return "-Infinity"
    return x


def world_to_js(w, colored=False):
    """An XpbdWorld-shaped dict (xpbd.js:188-240 field names) of a World, for
    the Node gpuworld surface; +-Infinity / NaN as strings."""
    cons = []
    for k in range(w.nc):
        cons.append(dict(type=int(w.type[k]), a=int(w.a[k]), b=int(w.b[k]), owner=int(w.owner[k]),
                         lax=_js_num(w.lax[k]), lay=_js_num(w.lay[k]), lbx=_js_num(w.lbx[k]), lby=_js_num(w.lby[k]),
                         compliance=_js_num(w.compliance[k]), rest=_js_num(w.rest[k]),
                         minAngle=_js_num(w.minA[k]), maxAngle=_js_num(w.maxA[k]), target=_js_num(w.target[k]),
                         breakForce=_js_num(w.breakForce[k]), broken=bool(w.broken[k]),
                         **{"lambda": _js_num(w.lambda_[k])}, lambdaX=_js_num(w.lambdaX[k]),
                         lambdaY=_js_num(w.lambdaY[k]), force=_js_num(w.force[k]),
                         color=int(w.color[k])))
    s = w.solver
    out = dict(n=int(w.n), substeps=int(s["substeps"]), iterations=int(s["iterations"]), h=float(s["h"]),
               linDrag=float(s["lin_drag"]), angDrag=float(s["ang_drag"]), maxSpeed=float(s["max_speed"]),
               maxAngSpeed=float(s["max_ang_speed"]),
               px=[float(v) for v in w.px], py=[float(v) for v in w.py], vx=[float(v) for v in w.vx],
               vy=[float(v) for v in w.vy], ang=[float(v) for v in w.ang], av=[float(v) for v in w.av],
               im=[float(v) for v in w.im], ii=[float(v) for v in w.ii], constraints=cons,
               collPts=[dict(body=int(b), lx=float(x), ly=float(y), r=float(r))
                        for b, x, y, r in zip(w.p_body, w.p_lx, w.p_ly, w.p_r)],
               rects=[dict(x=float(r[0]), y=float(r[1]), w=float(r[2]), h=float(r[3])) for r in w.rects])
    if colored and w.colorOff is not None:
        out["colorOff"], out["colorCnt"] = [int(v) for v in w.colorOff], [int(v) for v in w.colorCnt]
    return out


def creatures_to_js(w):
    """kx_creature_desc arm tables of a World (desc indices)."""
    return dict(nCreatures=int(w.n_creatures), nArms=int(len(w.arm_tip)),
                creatureArmOff=[int(v) for v in w.arm_off], armTip=[int(v) for v in w.arm_tip],
                armGrip=[int(v) for v in w.arm_grip], armMotorOff=[int(v) for v in w.motor_off],
                motors=[int(v) for v in w.motors])
