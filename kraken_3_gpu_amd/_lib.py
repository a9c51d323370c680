"""ctypes binding of libkxpbd.so (include/kxpbd.h).

The library is built in-tree by kraken_3_gpu_amd/build.py. Loading fails
loudly when it is missing: there is no CPU fallback for any entry point.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkxpbd.so")

KX_OK, KX_EINVAL, KX_ENOMEM, KX_EDEVICE, KX_ENOTSUP = 0, -22, -12, -5, -95
KX_PIN, KX_ANGLE, KX_GRIP = 0, 1, 2
KX_LAW_FIXED_GRIP = 1
KX_LAW_TAPE = 2
KX_TAPE_WORDS = 25
KX_PATH_PLAIN, KX_PATH_PERSISTENT = 0, 1
KX_FOLD_AUTO, KX_FOLD_PARALLEL, KX_FOLD_OWNER = 0, 1, 2
KX_STATE_EXTRA = 12

i32, u32, f64, f32p, i32p, u8p, f64p, u32p = (C.c_int32, C.c_uint32, C.c_double, C.POINTER(C.c_float),
                                              C.POINTER(C.c_int32), C.POINTER(C.c_uint8),
                                              C.POINTER(C.c_double), C.POINTER(C.c_uint32))


class WorldDesc(C.Structure):
    _fields_ = [("n_bodies", i32), ("n_constraints", i32), ("n_points", i32), ("n_rects", i32),
                ("inv_mass", f32p), ("inv_inertia", f32p),
                ("c_type", i32p), ("c_a", i32p), ("c_b", i32p), ("c_owner", i32p),
                ("c_lax", f32p), ("c_lay", f32p), ("c_lbx", f32p), ("c_lby", f32p),
                ("c_rest", f32p), ("c_min", f32p), ("c_max", f32p),
                ("c_color", i32p),
                ("p_body", i32p), ("p_lx", f32p), ("p_ly", f32p), ("p_r", f32p),
                ("rects", f32p)]


class CreatureDesc(C.Structure):
    _fields_ = [("n_creatures", i32), ("n_arms", i32), ("creature_arm_off", i32p), ("arm_tip", i32p),
                ("arm_grip", i32p), ("arm_motor_off", i32p), ("motors", i32p)]


class Params(C.Structure):
    _fields_ = [("substeps", i32), ("iterations", i32), ("h", f64), ("lin_drag", f64), ("ang_drag", f64),
                ("max_speed", f64), ("max_ang_speed", f64), ("path", i32), ("lanes_per_creature", i32),
                ("waves_per_block", i32)]


class BodyState(C.Structure):
    _fields_ = [(n, f32p) for n in ("px", "py", "vx", "vy", "ang", "av", "ppx", "ppy", "angP", "fx", "fy", "tq")]


class ConsState(C.Structure):
    _fields_ = [("lbx", f32p), ("lby", f32p), ("compliance", f32p), ("target", f32p), ("break_force", f32p),
                ("broken", u8p), ("lambda_", f32p), ("lambda_x", f32p), ("lambda_y", f32p)]


class Stats(C.Structure):
    _fields_ = [("max_speed", f64), ("max_ang_speed", f64), ("nan", C.c_int64)]


class Law(C.Structure):
    _fields_ = [("seed", u32), ("tone", f64), ("shoulder_limit", f64), ("joint_limit", f64),
                ("stiff_min", f64), ("stiff_max", f64), ("grip_min", f64), ("grip_max", f64),
                ("slip_max", f64), ("cooldown", f64), ("attach_min", f64), ("release_below", f64),
                ("creature_base", u32), ("flags", u32)]


class Request(C.Structure):
    _fields_ = [("arms", i32), ("segs", i32), ("ring", i32), ("pairs", i32), ("steps", i32),
                ("x", f64), ("y", f64), ("angle", f64), ("policy_seed", u32), ("noise_seed", u32),
                ("sigma", f64), ("tag", C.c_int64), ("fitness", f32p), ("final_state", f32p),
                ("final_cap", C.c_int64)]


class RequestDone(C.Structure):
    _fields_ = [("tag", C.c_int64), ("status", i32), ("n_creatures", i32), ("n_bodies", i32),
                ("t_submit", f64), ("t_admit", f64), ("t_done", f64), ("device_bytes", C.c_int64)]


class ServerOpts(C.Structure):
    _fields_ = [("chunk_steps", i32), ("max_active", i32), ("host_threads", i32), ("hbm_reserve", C.c_int64),
                ("params", Params), ("law", Law), ("map_width", f64), ("map_height", f64),
                ("map_density", f64), ("map_wall", f64), ("map_cols", i32), ("map_rows", i32),
                ("map_seed", u32)]


class ServerInfo(C.Structure):
    _fields_ = [("queued", i32), ("active", i32), ("done", i32), ("launches", C.c_int64),
                ("creature_steps", C.c_int64), ("now", f64), ("busy", f64), ("peak_active", C.c_int64),
                ("admit_waits", C.c_int64)]


class EsOpts(C.Structure):
    _fields_ = [("sigma", f64), ("lr", f64), ("min_sigma", f64), ("sigma_decay", f64), ("beta1", f64),
                ("beta2", f64), ("adam_eps", f64), ("clip", f64), ("pop_size", i32)]


EXPORTS = {
    "kx_create": (i32, [C.POINTER(WorldDesc), C.POINTER(CreatureDesc), C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "kx_destroy": (None, [C.c_void_p]),
    "kx_last_error": (C.c_char_p, []),
    "kx_abi_version": (i32, []),
    "kx_counts": (i32, [C.c_void_p, i32p, i32p, i32p, i32p]),
    "kx_constraint_order": (i32, [C.c_void_p, i32p, i32p]),
    "kx_color_info": (i32, [C.c_void_p, u32p, u32p]),
    "kx_group_points": (i32, [i32, i32p, i32, i32p, i32p]),
    "kx_color_constraints": (i32, [i32, i32, i32p, i32p, i32p, i32p, i32p]),
    "kx_plan_rounds": (i32, [i32, i32, i32p, i32p, i32p, i32, i32p, i32p, i32]),
    "kx_upload": (i32, [C.c_void_p, C.POINTER(BodyState), C.POINTER(ConsState)]),
    "kx_pack_controls": (i32, [C.c_void_p, C.POINTER(ConsState)]),
    "kx_download": (i32, [C.c_void_p, C.POINTER(BodyState), C.POINTER(ConsState), C.POINTER(Stats)]),
    "kx_get_cooldown": (i32, [C.c_void_p, f64p]),
    "kx_set_cooldown": (i32, [C.c_void_p, f64p]),
    "kx_step": (i32, [C.c_void_p, i32]),
    "kx_substep": (i32, [C.c_void_p, i32]),
    "kx_rollout": (i32, [C.c_void_p, C.POINTER(Law), i32, i32, u32p]),
    "kx_sync": (i32, [C.c_void_p]),
    "kx_rollout_noalloc": (i32, [C.c_void_p, C.POINTER(Law), i32, i32, u32p]),
    "kx_upload_tape": (i32, [C.c_void_p, C.POINTER(C.c_float), i32]),
    "kx_server_create": (i32, [C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    "kx_server_submit": (i32, [C.c_void_p, C.POINTER(Request)]),
    "kx_server_poll": (i32, [C.c_void_p, C.POINTER(RequestDone), i32, C.POINTER(i32), f64]),
    "kx_server_info_get": (i32, [C.c_void_p, C.POINTER(ServerInfo)]),
    "kx_server_destroy": (None, [C.c_void_p]),
    "kx_policy_tape": (i32, [C.POINTER(Request), i32, i32, f64, C.POINTER(C.c_float)]),
    "kx_request_bytes": (C.c_int64, [C.POINTER(Request), i32]),
    "kx_mem_info": (i32, [C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "kx_rollout_graph": (i32, [C.c_void_p, C.POINTER(Law), i32, i32, i32]),
    "kx_last_timing": (i32, [C.c_void_p, f64p, i32p]),
    "kx_set_path": (i32, [C.c_void_p, i32, i32, i32]),
    "kx_schedule_info": (i32, [C.c_void_p, i32p, i32]),
    "kx_persist_info": (i32, [C.c_void_p, i32p, i32]),
    "kx_set_register_slots": (i32, [C.c_void_p, i32]),
    "kx_set_collision_fold": (i32, [C.c_void_p, i32]),
    "kx_read_state": (i32, [C.c_void_p, f32p, i32]),
    "kx_respawn": (i32, [C.c_void_p, i32, i32p, f64p]),
    "kx_template_info": (i32, [C.c_void_p, i32, i32p]),
    "kx_set_template_lanes": (i32, [C.c_void_p, i32, i32p]),
    "kx_creature_templates": (i32, [C.c_void_p, i32p]),
    "kx_set_plain_sdf": (i32, [C.c_void_p, i32]),
    "kx_policy_weights": (i32, [i32, i32, i32]),
    "kx_policy_create": (i32, [i32, i32, i32, i32, C.POINTER(C.c_void_p)]),
    "kx_policy_destroy": (i32, [C.c_void_p]),
    "kx_policy_set_genomes": (i32, [C.c_void_p, f32p]),
    "kx_policy_get_genomes": (i32, [C.c_void_p, f32p]),
    "kx_policy_forward": (i32, [C.c_void_p, f32p, f32p, f32p, f32p, i32]),
    "kx_policy_last_ms": (i32, [C.c_void_p, f32p]),
    "kx_policy_drive": (i32, [C.c_void_p, C.c_void_p, f32p, i32, C.POINTER(Law), i32]),
    "kx_policy_forward_repeat": (i32, [C.c_void_p, i32, f32p]),
    "kx_es_create": (i32, [i32, C.POINTER(EsOpts), f32p, C.POINTER(C.c_void_p)]),
    "kx_es_destroy": (i32, [C.c_void_p]),
    "kx_es_generation": (i32, [C.c_void_p, C.c_void_p, u32p]),
    "kx_es_update": (i32, [C.c_void_p, f64p, f64p]),
    "kx_es_get": (i32, [C.c_void_p, f32p, f32p, f32p, f32p, f64p]),
    "kx_builder_new": (C.c_void_p, []),
    "kx_builder_free": (None, [C.c_void_p]),
    "kx_builder_map": (i32, [C.c_void_p, f64, f64, i32, i32, f64, f64, u32]),
    "kx_builder_find_spawn": (i32, [C.c_void_p, f64, u32p, f64p]),
    "kx_builder_crn_spawns": (i32, [C.c_void_p, i32, i32, u32p, f64p]),
    "kx_builder_spawn": (i32, [C.c_void_p, f64, f64, f64, i32, i32, i32, f64]),
    "kx_builder_desc": (i32, [C.c_void_p, C.POINTER(WorldDesc), C.POINTER(CreatureDesc)]),
    "kx_builder_state": (i32, [C.c_void_p, f64p, f64p, f64p]),
    "kx_builder_cons_init": (i32, [C.c_void_p, f64p, f64p, f64p, f64p, f64p, u8p]),
}

_lib = None


class KxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kxpbd error {code}: {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run kraken_3_gpu_amd/build.py (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise KxError(rc, lib().kx_last_error().decode(errors="replace"))
    return rc


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))
