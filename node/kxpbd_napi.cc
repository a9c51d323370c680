// Node-API binding of libkxpbd (include/kxpbd.h) for the gpuworld.js-shaped
// module in this directory. Plain C ABI calls on typed arrays; the two calls
// the reference makes asynchronous (step, download — src/gpu/gpuworld.js
// 374-396) return Promises and run on a worker thread (napi_async_work), the
// rest are synchronous. A negative return code becomes a thrown Error with
// the library's message, as gpuworld.js throws.
#include <node_api.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <deque>
#include <unordered_set>
#include <vector>

#include "../include/kxpbd.h"

namespace {

#define NAPI_OK(call)                                              \
  do {                                                             \
    if ((call) != napi_ok) {                                       \
      napi_throw_error(env, nullptr, "kxpbd: N-API call failed");  \
      return nullptr;                                              \
    }                                                              \
  } while (0)

napi_value throw_kx(napi_env env, int rc) {
  std::string m = "kxpbd error " + std::to_string(rc) + ": ";
  const char *e = kx_last_error();
  m += e ? e : "";
  napi_throw_error(env, nullptr, m.c_str());
  return nullptr;
}

bool get_prop(napi_env env, napi_value obj, const char *name, napi_value *out) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return false;
  if (napi_get_named_property(env, obj, name, out) != napi_ok) return false;
  napi_valuetype t;
  napi_typeof(env, *out, &t);
  return t != napi_undefined && t != napi_null;
}

// Typed-array view of obj[name] with the expected element type, or nullptr
// if absent; throws (returns false) on a wrong type or short array.
template <class T>
bool typed(napi_env env, napi_value obj, const char *name, napi_typedarray_type want, size_t min_len, T **out,
           bool required) {
  *out = nullptr;
  napi_value v;
  if (!get_prop(env, obj, name, &v)) {
    if (required) {
      napi_throw_type_error(env, nullptr, (std::string("kxpbd: missing typed array ") + name).c_str());
      return false;
    }
    return true;
  }
  bool is = false;
  napi_is_typedarray(env, v, &is);
  if (!is) {
    napi_throw_type_error(env, nullptr, (std::string("kxpbd: not a typed array: ") + name).c_str());
    return false;
  }
  napi_typedarray_type t;
  size_t len;
  void *data;
  napi_value ab;
  size_t off;
  napi_get_typedarray_info(env, v, &t, &len, &data, &ab, &off);
  if (t != want || len < min_len) {
    napi_throw_type_error(env, nullptr, (std::string("kxpbd: wrong type or length: ") + name).c_str());
    return false;
  }
  *out = static_cast<T *>(data);
  return true;
}

bool get_i32(napi_env env, napi_value obj, const char *name, int32_t *out, int32_t dflt) {
  napi_value v;
  *out = dflt;
  if (!get_prop(env, obj, name, &v)) return true;
  return napi_get_value_int32(env, v, out) == napi_ok;
}

bool get_f64(napi_env env, napi_value obj, const char *name, double *out, double dflt) {
  napi_value v;
  *out = dflt;
  if (!get_prop(env, obj, name, &v)) return true;
  return napi_get_value_double(env, v, out) == napi_ok;
}

// A world handle's external. Asynchronous calls on one handle run one at a
// time in call order (the WebGPU queue's FIFO of the reference); a destroy
// while any is queued or running takes effect when the last one completes,
// and synchronous calls are refused meanwhile instead of touching a handle a
// worker thread is using.
struct Work;
struct WorldSlot {
  kx_handle *h = nullptr;
  std::deque<Work *> pending;   // queued, not started
  bool running = false;         // one Work on a worker thread
  bool destroy_pending = false;
};

WorldSlot *slot_of(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_error(env, nullptr, "kxpbd: not a live handle");
    return nullptr;
  }
  WorldSlot *s = static_cast<WorldSlot *>(p);
  if (!s->h || s->destroy_pending) {
    napi_throw_error(env, nullptr, "kxpbd: not a live handle");
    return nullptr;
  }
  return s;
}

// the handle for an asynchronous call (queued behind the handle's work)
kx_handle *handle_of(napi_env env, napi_value v) {
  WorldSlot *s = slot_of(env, v);
  return s ? s->h : nullptr;
}

// the handle for a synchronous call: refused while asynchronous work on it
// is queued or running
kx_handle *idle_handle_of(napi_env env, napi_value v) {
  WorldSlot *s = slot_of(env, v);
  if (!s) return nullptr;
  if (s->running || !s->pending.empty()) {
    napi_throw_error(env, nullptr, "kxpbd: handle busy: await its pending step/download/rollout first");
    return nullptr;
  }
  return s->h;
}

void finalize_handle(napi_env, void *data, void *) {
  // the external stays reachable from every queued Work, so no work is left
  WorldSlot *s = static_cast<WorldSlot *>(data);
  if (s->h) kx_destroy(s->h);
  delete s;
}

// create(desc, params[, creatures]) -> handle; `creatures` (optional) carries
// the arm tables of kx_creature_desc for the device-side control law
napi_value Create(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 2) {
    napi_throw_type_error(env, nullptr, "kxpbd.create(desc, params)");
    return nullptr;
  }
  kx_world_desc d{};
  int32_t nb, nc, np, nr;
  if (!get_i32(env, a[0], "nBodies", &nb, 0) || !get_i32(env, a[0], "nConstraints", &nc, 0) ||
      !get_i32(env, a[0], "nPoints", &np, 0) || !get_i32(env, a[0], "nRects", &nr, 0)) {
    napi_throw_type_error(env, nullptr, "kxpbd: bad counts");
    return nullptr;
  }
  d.n_bodies = nb; d.n_constraints = nc; d.n_points = np; d.n_rects = nr;
  const size_t B = nb > 0 ? nb : 0, Cn = nc > 0 ? nc : 0, P = np > 0 ? np : 0, R = nr > 0 ? nr : 0;
  float *fp;
  int32_t *ip;
#define F32(field, key, n, req) \
  if (!typed(env, a[0], key, napi_float32_array, n, &fp, req)) return nullptr; \
  d.field = fp;
#define I32(field, key, n, req) \
  if (!typed(env, a[0], key, napi_int32_array, n, &ip, req)) return nullptr; \
  d.field = ip;
  F32(inv_mass, "invMass", B, true);
  F32(inv_inertia, "invInertia", B, true);
  I32(c_type, "cType", Cn, true);
  I32(c_a, "cA", Cn, true);
  I32(c_b, "cB", Cn, true);
  I32(c_owner, "cOwner", Cn, false);
  F32(c_lax, "cLax", Cn, true);
  F32(c_lay, "cLay", Cn, true);
  F32(c_lbx, "cLbx", Cn, true);
  F32(c_lby, "cLby", Cn, true);
  F32(c_rest, "cRest", Cn, true);
  F32(c_min, "cMin", Cn, true);
  F32(c_max, "cMax", Cn, true);
  I32(c_color, "cColor", Cn, false);
  I32(p_body, "pBody", P, true);
  F32(p_lx, "pLx", P, true);
  F32(p_ly, "pLy", P, true);
  F32(p_r, "pR", P, true);
  F32(rects, "rects", 4 * R, true);
#undef F32
#undef I32
  kx_params p{};
  double v;
  get_i32(env, a[1], "substeps", &p.substeps, 4);
  get_i32(env, a[1], "iterations", &p.iterations, 8);
  get_f64(env, a[1], "h", &v, 1.0 / 240.0); p.h = v;
  get_f64(env, a[1], "linDrag", &v, 3.0); p.lin_drag = v;
  get_f64(env, a[1], "angDrag", &v, 3.0); p.ang_drag = v;
  get_f64(env, a[1], "maxSpeed", &v, 360.0); p.max_speed = v;
  get_f64(env, a[1], "maxAngSpeed", &v, 18.0); p.max_ang_speed = v;
  get_i32(env, a[1], "path", &p.path, KX_PATH_PERSISTENT);
  get_i32(env, a[1], "lanesPerCreature", &p.lanes_per_creature, 0);
  get_i32(env, a[1], "wavesPerBlock", &p.waves_per_block, 0);
  kx_creature_desc cr{};
  bool has_cr = false;
  if (argc >= 3) {
    napi_valuetype t;
    napi_typeof(env, a[2], &t);
    if (t == napi_object) {
      int32_t ncr = 0, narm = 0;
      get_i32(env, a[2], "nCreatures", &ncr, 0);
      get_i32(env, a[2], "nArms", &narm, 0);
      cr.n_creatures = ncr;
      cr.n_arms = narm;
      int32_t *q;
      const size_t NC1 = ncr >= 0 ? (size_t)ncr + 1 : 0, NA = narm >= 0 ? (size_t)narm : 0;
      if (!typed(env, a[2], "creatureArmOff", napi_int32_array, NC1, &q, true)) return nullptr;
      cr.creature_arm_off = q;
      if (!typed(env, a[2], "armTip", napi_int32_array, NA, &q, true)) return nullptr;
      cr.arm_tip = q;
      if (!typed(env, a[2], "armGrip", napi_int32_array, NA, &q, true)) return nullptr;
      cr.arm_grip = q;
      if (!typed(env, a[2], "armMotorOff", napi_int32_array, NA + 1, &q, true)) return nullptr;
      cr.arm_motor_off = q;
      const int32_t nmot = q[NA];
      if (nmot < 0) {
        napi_throw_range_error(env, nullptr, "kxpbd: armMotorOff[nArms] < 0");
        return nullptr;
      }
      if (!typed(env, a[2], "motors", napi_int32_array, (size_t)nmot, &q, true)) return nullptr;
      cr.motors = q;
      has_cr = true;
    }
  }
  kx_handle *h = nullptr;
  int rc = kx_create(&d, has_cr ? &cr : nullptr, &p, &h);
  if (rc < 0) return throw_kx(env, rc);
  WorldSlot *slot = new WorldSlot();
  slot->h = h;
  napi_value ext;
  if (napi_create_external(env, slot, finalize_handle, nullptr, &ext) != napi_ok) {
    kx_destroy(h);
    delete slot;
    napi_throw_error(env, nullptr, "kxpbd: napi_create_external failed");
    return nullptr;
  }
  return ext;
}

// destroy(handle): idempotent; with asynchronous work queued or running the
// release happens when the last of it completes (the handle is dead to new
// calls at once)
napi_value Destroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  void *p = nullptr;
  if (argc >= 1 && napi_get_value_external(env, a[0], &p) == napi_ok && p) {
    WorldSlot *slot = static_cast<WorldSlot *>(p);
    if (slot->running || !slot->pending.empty()) {
      slot->destroy_pending = true;
    } else if (slot->h) {
      kx_destroy(slot->h);
      slot->h = nullptr;
    }
  }
  return nullptr;
}

napi_value Counts(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  kx_handle *h = idle_handle_of(env, a[0]);
  if (!h) return nullptr;
  int32_t c[4];
  int rc = kx_counts(h, &c[0], &c[1], &c[2], &c[3]);
  if (rc < 0) return throw_kx(env, rc);
  napi_value arr;
  NAPI_OK(napi_create_array_with_length(env, 4, &arr));
  for (uint32_t i = 0; i < 4; i++) {
    napi_value x;
    NAPI_OK(napi_create_int32(env, c[i], &x));
    NAPI_OK(napi_set_element(env, arr, i, x));
  }
  return arr;
}

// constraintOrder(handle, {order:Int32Array, color:Int32Array})
napi_value ConstraintOrder(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  kx_handle *h = idle_handle_of(env, a[0]);
  if (!h) return nullptr;
  int32_t nb, nc, np, ncr;
  int rc = kx_counts(h, &nb, &nc, &np, &ncr);
  if (rc < 0) return throw_kx(env, rc);
  int32_t *order, *color;
  if (!typed(env, a[1], "order", napi_int32_array, nc, &order, true) ||
      !typed(env, a[1], "color", napi_int32_array, nc, &color, true))
    return nullptr;
  rc = kx_constraint_order(h, order, color);
  if (rc < 0) return throw_kx(env, rc);
  return nullptr;
}

// colorInfo(handle, {off:Uint32Array(16), cnt:Uint32Array(16)})
napi_value ColorInfo(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  kx_handle *h = idle_handle_of(env, a[0]);
  if (!h) return nullptr;
  uint32_t *off, *cnt;
  if (!typed(env, a[1], "off", napi_uint32_array, 16, &off, true) ||
      !typed(env, a[1], "cnt", napi_uint32_array, 16, &cnt, true))
    return nullptr;
  int rc = kx_color_info(h, off, cnt);
  if (rc < 0) return throw_kx(env, rc);
  return nullptr;
}

bool body_state(napi_env env, napi_value o, size_t n, kx_body_state *b, bool need_all) {
  *b = kx_body_state{};
  return typed(env, o, "px", napi_float32_array, n, &b->px, need_all) &&
         typed(env, o, "py", napi_float32_array, n, &b->py, need_all) &&
         typed(env, o, "vx", napi_float32_array, n, &b->vx, need_all) &&
         typed(env, o, "vy", napi_float32_array, n, &b->vy, need_all) &&
         typed(env, o, "ang", napi_float32_array, n, &b->ang, need_all) &&
         typed(env, o, "av", napi_float32_array, n, &b->av, need_all) &&
         typed(env, o, "ppx", napi_float32_array, n, &b->ppx, false) &&
         typed(env, o, "ppy", napi_float32_array, n, &b->ppy, false) &&
         typed(env, o, "angP", napi_float32_array, n, &b->angP, false) &&
         typed(env, o, "fx", napi_float32_array, n, &b->fx, false) &&
         typed(env, o, "fy", napi_float32_array, n, &b->fy, false) &&
         typed(env, o, "tq", napi_float32_array, n, &b->tq, false);
}

bool cons_state(napi_env env, napi_value o, size_t n, kx_cons_state *c) {
  *c = kx_cons_state{};
  return typed(env, o, "lbx", napi_float32_array, n, &c->lbx, false) &&
         typed(env, o, "lby", napi_float32_array, n, &c->lby, false) &&
         typed(env, o, "compliance", napi_float32_array, n, &c->compliance, false) &&
         typed(env, o, "target", napi_float32_array, n, &c->target, false) &&
         typed(env, o, "breakForce", napi_float32_array, n, &c->break_force, false) &&
         typed(env, o, "broken", napi_uint8_array, n, &c->broken, false) &&
         typed(env, o, "lambda", napi_float32_array, n, &c->lambda, false) &&
         typed(env, o, "lambdaX", napi_float32_array, n, &c->lambda_x, false) &&
         typed(env, o, "lambdaY", napi_float32_array, n, &c->lambda_y, false);
}

// upload(handle, bodies, cons)
napi_value Upload(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  kx_handle *h = idle_handle_of(env, a[0]);
  if (!h) return nullptr;
  int32_t nb, nc, np, ncr;
  int rc = kx_counts(h, &nb, &nc, &np, &ncr);
  if (rc < 0) return throw_kx(env, rc);
  kx_body_state b;
  kx_cons_state c;
  if (!body_state(env, a[1], nb, &b, true) || !cons_state(env, a[2], nc, &c)) return nullptr;
  rc = kx_upload(h, &b, &c);
  if (rc < 0) return throw_kx(env, rc);
  return nullptr;
}

// packControls(handle, cons)
napi_value PackControls(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  kx_handle *h = idle_handle_of(env, a[0]);
  if (!h) return nullptr;
  int32_t nb, nc, np, ncr;
  int rc = kx_counts(h, &nb, &nc, &np, &ncr);
  if (rc < 0) return throw_kx(env, rc);
  kx_cons_state c;
  if (!cons_state(env, a[1], nc, &c)) return nullptr;
  rc = kx_pack_controls(h, &c);
  if (rc < 0) return throw_kx(env, rc);
  return nullptr;
}

// ── asynchronous calls ──────────────────────────────────────────────
struct Work {
  napi_async_work work = nullptr;
  napi_deferred deferred = nullptr;
  std::vector<napi_ref> keep;         // the handle and typed arrays, alive until completion
  WorldSlot *slot = nullptr;
  kx_handle *h = nullptr;
  int kind = 0;                       // 0 step, 1 download, 2 rollout, 3 substep
  int32_t nsteps = 0, step0 = 0;
  kx_body_state b{};
  kx_cons_state c{};
  kx_stats st{};
  kx_law law{};
  uint32_t *events = nullptr;
  int rc = 0;
  std::string err;
};

void exec_work(napi_env, void *data) {
  Work *w = static_cast<Work *>(data);
  if (w->kind == 0 || w->kind == 3) {
    w->rc = w->kind == 0 ? kx_step(w->h, w->nsteps) : kx_substep(w->h, w->nsteps);
    if (w->rc >= 0) w->rc = kx_sync(w->h);
  } else if (w->kind == 1) {
    w->rc = kx_download(w->h, &w->b, &w->c, &w->st);
  } else {
    w->rc = kx_rollout(w->h, &w->law, w->step0, w->nsteps, w->events);
    if (w->rc >= 0) w->rc = kx_sync(w->h);
  }
  if (w->rc < 0) {
    const char *e = kx_last_error();   // thread-local: read on this thread
    w->err = "kxpbd error " + std::to_string(w->rc) + ": " + (e ? e : "");
  }
}

// starts the handle's next queued call when none is running; releases a
// destroyed handle once its work has drained
void start_next(napi_env env, WorldSlot *slot) {
  if (slot->running) return;
  if (!slot->pending.empty()) {
    Work *w = slot->pending.front();
    slot->pending.pop_front();
    slot->running = true;
    napi_queue_async_work(env, w->work);
    return;
  }
  if (slot->destroy_pending && slot->h) {
    kx_destroy(slot->h);
    slot->h = nullptr;
  }
}

void done_work(napi_env env, napi_status, void *data) {
  Work *w = static_cast<Work *>(data);
  napi_value result = nullptr;
  if (w->rc < 0) {
    napi_value msg;
    napi_create_string_utf8(env, w->err.c_str(), NAPI_AUTO_LENGTH, &msg);
    napi_create_error(env, nullptr, msg, &result);
    napi_reject_deferred(env, w->deferred, result);
  } else {
    if (w->kind == 1) {
      napi_create_object(env, &result);
      napi_value x;
      napi_create_double(env, w->st.max_speed, &x);
      napi_set_named_property(env, result, "maxSpeed", x);
      napi_create_double(env, w->st.max_ang_speed, &x);
      napi_set_named_property(env, result, "maxAngSpeed", x);
      napi_create_double(env, (double)w->st.nan, &x);
      napi_set_named_property(env, result, "nan", x);
    } else {
      napi_get_undefined(env, &result);
    }
    napi_resolve_deferred(env, w->deferred, result);
  }
  WorldSlot *slot = w->slot;
  for (napi_ref r : w->keep) napi_delete_reference(env, r);
  napi_delete_async_work(env, w->work);
  delete w;
  slot->running = false;
  start_next(env, slot);
}

// `keep[0]` is the handle's external
napi_value queue(napi_env env, Work *w, const std::vector<napi_value> &keep) {
  napi_value promise, name;
  void *p = nullptr;
  if (napi_get_value_external(env, keep[0], &p) != napi_ok || !p) {
    delete w;
    napi_throw_error(env, nullptr, "kxpbd: not a live handle");
    return nullptr;
  }
  if (napi_create_promise(env, &w->deferred, &promise) != napi_ok) {
    delete w;
    napi_throw_error(env, nullptr, "kxpbd: napi_create_promise failed");
    return nullptr;
  }
  w->slot = static_cast<WorldSlot *>(p);
  for (napi_value v : keep) {
    napi_ref r;
    napi_create_reference(env, v, 1, &r);
    w->keep.push_back(r);
  }
  napi_create_string_utf8(env, "kxpbd", NAPI_AUTO_LENGTH, &name);
  napi_create_async_work(env, nullptr, name, exec_work, done_work, w, &w->work);
  w->slot->pending.push_back(w);
  start_next(env, w->slot);
  return promise;
}

// step(handle, nsteps) -> Promise<void>
napi_value Step(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  kx_handle *h = handle_of(env, a[0]);
  if (!h) return nullptr;
  int32_t n = 1;
  if (argc >= 2) napi_get_value_int32(env, a[1], &n);
  Work *w = new Work;
  w->h = h; w->kind = 0; w->nsteps = n;
  return queue(env, w, {a[0]});
}

// substep(handle, n) -> Promise<void> (gpuworld.js:344-372: no grip break)
napi_value Substep(napi_env env, napi_callback_info info) {
  size_t argc = 2;
  napi_value a[2];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  kx_handle *h = handle_of(env, a[0]);
  if (!h) return nullptr;
  int32_t n = 1;
  if (argc >= 2) napi_get_value_int32(env, a[1], &n);
  Work *w = new Work;
  w->h = h; w->kind = 3; w->nsteps = n;
  return queue(env, w, {a[0]});
}

// respawn(handle, creatures:Int32Array, poses:Float64Array(3n)) — in place on
// the device (kx_respawn; train.js:188-229 respawnCreature)
napi_value Respawn(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 3) {
    napi_throw_type_error(env, nullptr, "kxpbd.respawn(handle, creatures, poses)");
    return nullptr;
  }
  kx_handle *h = idle_handle_of(env, a[0]);
  if (!h) return nullptr;
  napi_typedarray_type t1, t2;
  size_t n1, n2, off;
  void *d1, *d2;
  napi_value ab;
  bool is1 = false, is2 = false;
  napi_is_typedarray(env, a[1], &is1);
  napi_is_typedarray(env, a[2], &is2);
  if (!is1 || !is2 || napi_get_typedarray_info(env, a[1], &t1, &n1, &d1, &ab, &off) != napi_ok ||
      napi_get_typedarray_info(env, a[2], &t2, &n2, &d2, &ab, &off) != napi_ok || t1 != napi_int32_array ||
      t2 != napi_float64_array || n2 < 3 * n1) {
    napi_throw_type_error(env, nullptr, "kxpbd.respawn: an Int32Array of creatures and a Float64Array of 3 per creature");
    return nullptr;
  }
  int rc = kx_respawn(h, (int32_t)n1, static_cast<int32_t *>(d1), static_cast<double *>(d2));
  if (rc < 0) return throw_kx(env, rc);
  return nullptr;
}

// download(handle, bodies, cons) -> Promise<{maxSpeed, maxAngSpeed, nan}>
napi_value Download(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  kx_handle *h = handle_of(env, a[0]);
  if (!h) return nullptr;
  int32_t nb, nc, np, ncr;
  int rc = kx_counts(h, &nb, &nc, &np, &ncr);
  if (rc < 0) return throw_kx(env, rc);
  Work *w = new Work;
  w->h = h; w->kind = 1;
  if (!body_state(env, a[1], nb, &w->b, false) || !cons_state(env, a[2], nc, &w->c)) {
    delete w;
    return nullptr;
  }
  return queue(env, w, {a[0], a[1], a[2]});
}

// readState(handle, {out:Float32Array}, stride) -> floats per creature written
// (kx_read_state: readProprio's inputs then the mantle state, per creature)
napi_value ReadState(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 3) {
    napi_throw_type_error(env, nullptr, "kxpbd.readState(handle, {out}, stride)");
    return nullptr;
  }
  kx_handle *h = idle_handle_of(env, a[0]);
  if (!h) return nullptr;
  int32_t stride = 0, nb, nc, np, ncr;
  NAPI_OK(napi_get_value_int32(env, a[2], &stride));
  int rc = kx_counts(h, &nb, &nc, &np, &ncr);
  if (rc < 0) return throw_kx(env, rc);
  if (stride <= 0) {
    napi_throw_range_error(env, nullptr, "kxpbd.readState: stride must be positive");
    return nullptr;
  }
  float *out;
  if (!typed(env, a[1], "out", napi_float32_array, (size_t)ncr * (size_t)stride, &out, true)) return nullptr;
  rc = kx_read_state(h, out, stride);
  if (rc < 0) return throw_kx(env, rc);
  napi_value r;
  NAPI_OK(napi_create_int32(env, rc, &r));
  return r;
}

// ── policy network and ES (kx_policy_*, kx_es_*) ──────────────────────
// Their externals hold a BrainSlot, registered in g_brain so a foreign
// external is refused before it is read; the leading null word makes
// handle_of refuse a BrainSlot passed where a world handle is expected.
enum { KIND_POLICY = 1, KIND_ES = 2 };
struct BrainSlot {
  void *world_guard = nullptr;
  int kind = 0;
  void *obj = nullptr;
  // shapes fixed at create: policy n, nIn, hidden, nOut, nWeights; ES
  // nWeights and the (even) population in dims[4], dims[5]
  int32_t dims[6] = {0, 0, 0, 0, 0, 0};
};
std::unordered_set<BrainSlot *> g_brain;

void finalize_brain(napi_env, void *data, void *) {
  BrainSlot *b = static_cast<BrainSlot *>(data);
  if (b->obj) {
    if (b->kind == KIND_POLICY) kx_policy_destroy(static_cast<kx_policy *>(b->obj));
    else kx_es_destroy(static_cast<kx_es *>(b->obj));
  }
  g_brain.erase(b);
  delete b;
}

BrainSlot *brain_slot(napi_env env, napi_value v, int kind) {
  void *p = nullptr;
  BrainSlot *b = nullptr;
  if (napi_get_value_external(env, v, &p) == napi_ok && p && g_brain.count(static_cast<BrainSlot *>(p)))
    b = static_cast<BrainSlot *>(p);
  if (!b || b->kind != kind || !b->obj) {
    napi_throw_error(env, nullptr, kind == KIND_POLICY ? "kxpbd: not a live policy" : "kxpbd: not a live ES");
    return nullptr;
  }
  return b;
}

void *brain_of(napi_env env, napi_value v, int kind) {
  BrainSlot *b = brain_slot(env, v, kind);
  return b ? b->obj : nullptr;
}

// a shape argument from JS must equal the object's own (a mismatch would
// size the typed-array checks wrongly)
bool same_dim(napi_env env, const char *what, int32_t given, int32_t own) {
  if (given == own) return true;
  std::string m = std::string("kxpbd: ") + what + " " + std::to_string(given) + " does not match the object's " +
                  std::to_string(own);
  napi_throw_range_error(env, nullptr, m.c_str());
  return false;
}

napi_value wrap_brain(napi_env env, int kind, void *obj, const int32_t *dims) {
  BrainSlot *b = new BrainSlot();
  b->kind = kind;
  b->obj = obj;
  for (int i = 0; i < 6; i++) b->dims[i] = dims[i];
  napi_value ext;
  if (napi_create_external(env, b, finalize_brain, nullptr, &ext) != napi_ok) {
    if (kind == KIND_POLICY) kx_policy_destroy(static_cast<kx_policy *>(obj));
    else kx_es_destroy(static_cast<kx_es *>(obj));
    delete b;
    napi_throw_error(env, nullptr, "kxpbd: napi_create_external failed");
    return nullptr;
  }
  g_brain.insert(b);
  return ext;
}

bool get_i32(napi_env env, napi_value v, int32_t *out) {
  if (napi_get_value_int32(env, v, out) != napi_ok) {
    napi_throw_type_error(env, nullptr, "kxpbd: expected an integer");
    return false;
  }
  return true;
}

bool get_f64_prop(napi_env env, napi_value obj, const char *name, double *out) {
  napi_value v;
  if (!get_prop(env, obj, name, &v)) return true;   // keep the default
  if (napi_get_value_double(env, v, out) != napi_ok) {
    napi_throw_type_error(env, nullptr, (std::string("kxpbd: not a number: ") + name).c_str());
    return false;
  }
  return true;
}

// policyCreate(n, nIn, hidden, nOut) -> policy (src/brain/net.js layout)
napi_value PolicyCreate(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 4) {
    napi_throw_type_error(env, nullptr, "kxpbd.policyCreate(n, nIn, hidden, nOut)");
    return nullptr;
  }
  int32_t v[4];
  for (int i = 0; i < 4; i++)
    if (!get_i32(env, a[i], &v[i])) return nullptr;
  kx_policy *p = nullptr;
  int rc = kx_policy_create(v[0], v[1], v[2], v[3], &p);
  if (rc < 0) return throw_kx(env, rc);
  const int32_t dims[6] = {v[0], v[1], v[2], v[3], kx_policy_weights(v[1], v[2], v[3]), 0};
  return wrap_brain(env, KIND_POLICY, p, dims);
}

// policyWeights(nIn, hidden, nOut) -> weightCount (net.js:30-32)
napi_value PolicyWeights(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  int32_t v[3] = {0, 24, 20};
  for (size_t i = 0; i < argc && i < 3; i++)
    if (!get_i32(env, a[i], &v[i])) return nullptr;
  napi_value r;
  NAPI_OK(napi_create_int32(env, kx_policy_weights(v[0], v[1], v[2]), &r));
  return r;
}

// policyGenomes(policy, {genomes:Float32Array}, upload:boolean)
napi_value PolicyGenomes(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value a[5];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 5) {
    napi_throw_type_error(env, nullptr, "kxpbd.policyGenomes(policy, {genomes}, upload, n, nWeights)");
    return nullptr;
  }
  BrainSlot *b = brain_slot(env, a[0], KIND_POLICY);
  if (!b) return nullptr;
  kx_policy *p = static_cast<kx_policy *>(b->obj);
  bool up = false;
  NAPI_OK(napi_get_value_bool(env, a[2], &up));
  int32_t n = 0, nw = 0;
  if (!get_i32(env, a[3], &n) || !get_i32(env, a[4], &nw)) return nullptr;
  if (!same_dim(env, "population", n, b->dims[0]) || !same_dim(env, "weights", nw, b->dims[4])) return nullptr;
  float *g;
  if (!typed(env, a[1], "genomes", napi_float32_array, (size_t)n * (size_t)nw, &g, true)) return nullptr;
  int rc = up ? kx_policy_set_genomes(p, g) : kx_policy_get_genomes(p, g);
  if (rc < 0) return throw_kx(env, rc);
  return nullptr;
}

// policyForward(policy, {x, out, scratch?, controls?}, arms, n, nIn, nOut,
// hidden) -> kernel milliseconds; lengths are checked against n and the shape
napi_value PolicyForward(napi_env env, napi_callback_info info) {
  size_t argc = 7;
  napi_value a[7];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 7) {
    napi_throw_type_error(env, nullptr, "kxpbd.policyForward(policy, {x, out, scratch, controls}, arms, n, nIn, nOut, hidden)");
    return nullptr;
  }
  BrainSlot *b = brain_slot(env, a[0], KIND_POLICY);
  if (!b) return nullptr;
  kx_policy *p = static_cast<kx_policy *>(b->obj);
  int32_t arms, n, n_in, n_out, hidden;
  if (!get_i32(env, a[2], &arms) || !get_i32(env, a[3], &n) || !get_i32(env, a[4], &n_in) ||
      !get_i32(env, a[5], &n_out) || !get_i32(env, a[6], &hidden))
    return nullptr;
  if (!same_dim(env, "population", n, b->dims[0]) || !same_dim(env, "inputs", n_in, b->dims[1]) ||
      !same_dim(env, "hidden", hidden, b->dims[2]) || !same_dim(env, "outputs", n_out, b->dims[3]))
    return nullptr;
  float *x, *out, *scratch, *controls;
  if (!typed(env, a[1], "x", napi_float32_array, (size_t)n * n_in, &x, true) ||
      !typed(env, a[1], "out", napi_float32_array, (size_t)n * n_out, &out, true) ||
      !typed(env, a[1], "scratch", napi_float32_array, (size_t)n * 2 * hidden, &scratch, false) ||
      !typed(env, a[1], "controls", napi_float32_array, (size_t)n * KX_TAPE_WORDS, &controls, false))
    return nullptr;
  int rc = kx_policy_forward(p, x, out, scratch, controls, arms);
  if (rc < 0) return throw_kx(env, rc);
  float ms = 0.f;
  kx_policy_last_ms(p, &ms);
  napi_value r;
  NAPI_OK(napi_create_double(env, ms, &r));
  return r;
}

// esCreate(nWeights, opts, {theta?}) -> es (src/brain/es.js createES option names)
napi_value EsCreate(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 2) {
    napi_throw_type_error(env, nullptr, "kxpbd.esCreate(nWeights, opts, {theta})");
    return nullptr;
  }
  int32_t nw;
  if (!get_i32(env, a[0], &nw)) return nullptr;
  kx_es_opts o{0.06, 0.030, 0.004, 0.999, 0.9, 0.999, 1e-8, 2.5, 64};   // es.js:23-38
  double pop = o.pop_size;
  if (!get_f64_prop(env, a[1], "sigma", &o.sigma) || !get_f64_prop(env, a[1], "lr", &o.lr) ||
      !get_f64_prop(env, a[1], "minSigma", &o.min_sigma) || !get_f64_prop(env, a[1], "sigmaDecay", &o.sigma_decay) ||
      !get_f64_prop(env, a[1], "adamBeta1", &o.beta1) || !get_f64_prop(env, a[1], "adamBeta2", &o.beta2) ||
      !get_f64_prop(env, a[1], "adamEps", &o.adam_eps) || !get_f64_prop(env, a[1], "clip", &o.clip) ||
      !get_f64_prop(env, a[1], "popSize", &pop))
    return nullptr;
  o.pop_size = (int32_t)pop;
  float *theta = nullptr;
  if (argc >= 3 && !typed(env, a[2], "theta", napi_float32_array, (size_t)nw, &theta, false)) return nullptr;
  kx_es *e = nullptr;
  int rc = kx_es_create(nw, &o, theta, &e);
  if (rc < 0) return throw_kx(env, rc);
  // the library keeps the population even (es.js:37)
  const int32_t dims[6] = {0, 0, 0, 0, nw, o.pop_size + (o.pop_size & 1)};
  return wrap_brain(env, KIND_ES, e, dims);
}

// esGeneration(es, policy, rndState) -> rndState after the draws
napi_value EsGeneration(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 3) {
    napi_throw_type_error(env, nullptr, "kxpbd.esGeneration(es, policy, rndState)");
    return nullptr;
  }
  kx_es *e = static_cast<kx_es *>(brain_of(env, a[0], KIND_ES));
  if (!e) return nullptr;
  kx_policy *p = static_cast<kx_policy *>(brain_of(env, a[1], KIND_POLICY));
  if (!p) return nullptr;
  uint32_t st = 0;
  NAPI_OK(napi_get_value_uint32(env, a[2], &st));
  int rc = kx_es_generation(e, p, &st);
  if (rc < 0) return throw_kx(env, rc);
  napi_value r;
  NAPI_OK(napi_create_uint32(env, st, &r));
  return r;
}

napi_value stats_object(napi_env env, const char *const *names, const double *v, int n) {
  napi_value o;
  NAPI_OK(napi_create_object(env, &o));
  for (int i = 0; i < n; i++) {
    napi_value x;
    NAPI_OK(napi_create_double(env, v[i], &x));
    NAPI_OK(napi_set_named_property(env, o, names[i], x));
  }
  return o;
}

// esUpdate(es, {fitness:Float64Array}, popSize) -> {best, mean, sd, sigma}
napi_value EsUpdate(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 3) {
    napi_throw_type_error(env, nullptr, "kxpbd.esUpdate(es, {fitness}, popSize)");
    return nullptr;
  }
  BrainSlot *b = brain_slot(env, a[0], KIND_ES);
  if (!b) return nullptr;
  kx_es *e = static_cast<kx_es *>(b->obj);
  int32_t pop;
  if (!get_i32(env, a[2], &pop)) return nullptr;
  if (!same_dim(env, "popSize", pop, b->dims[5])) return nullptr;
  double *fit;
  if (!typed(env, a[1], "fitness", napi_float64_array, (size_t)pop, &fit, true)) return nullptr;
  double st[4];
  int rc = kx_es_update(e, fit, st);
  if (rc < 0) return throw_kx(env, rc);
  static const char *const names[4] = {"best", "mean", "sd", "sigma"};
  return stats_object(env, names, st, 4);
}

// esGet(es, {theta?, m?, v?, bestTheta?}, nWeights) -> {sigma, t, bestEver}
napi_value EsGet(napi_env env, napi_callback_info info) {
  size_t argc = 3;
  napi_value a[3];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 3) {
    napi_throw_type_error(env, nullptr, "kxpbd.esGet(es, {theta, m, v, bestTheta}, nWeights)");
    return nullptr;
  }
  BrainSlot *b = brain_slot(env, a[0], KIND_ES);
  if (!b) return nullptr;
  kx_es *e = static_cast<kx_es *>(b->obj);
  int32_t nw;
  if (!get_i32(env, a[2], &nw)) return nullptr;
  if (!same_dim(env, "nWeights", nw, b->dims[4])) return nullptr;
  float *th, *m, *v, *best;
  if (!typed(env, a[1], "theta", napi_float32_array, (size_t)nw, &th, false) ||
      !typed(env, a[1], "m", napi_float32_array, (size_t)nw, &m, false) ||
      !typed(env, a[1], "v", napi_float32_array, (size_t)nw, &v, false) ||
      !typed(env, a[1], "bestTheta", napi_float32_array, (size_t)nw, &best, false))
    return nullptr;
  double sc[3];
  int rc = kx_es_get(e, th, m, v, best, sc);
  if (rc < 0) return throw_kx(env, rc);
  static const char *const names[3] = {"sigma", "t", "bestEver"};
  return stats_object(env, names, sc, 3);
}

// brainDestroy(policy or es): frees now; the finalizer then has nothing left
napi_value BrainDestroy(napi_env env, napi_callback_info info) {
  size_t argc = 1;
  napi_value a[1];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  void *p = nullptr;
  if (argc >= 1 && napi_get_value_external(env, a[0], &p) == napi_ok && p &&
      g_brain.count(static_cast<BrainSlot *>(p))) {
    BrainSlot *b = static_cast<BrainSlot *>(p);
    if (b->obj) {
      if (b->kind == KIND_POLICY) kx_policy_destroy(static_cast<kx_policy *>(b->obj));
      else kx_es_destroy(static_cast<kx_es *>(b->obj));
    }
    b->obj = nullptr;
  }
  return nullptr;
}

// rollout(handle, law, step0, nsteps, events?:Uint32Array) -> Promise<void>
napi_value Rollout(napi_env env, napi_callback_info info) {
  size_t argc = 5;
  napi_value a[5];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 4) {
    napi_throw_type_error(env, nullptr, "kxpbd.rollout(handle, law, step0, nsteps[, events])");
    return nullptr;
  }
  kx_handle *h = handle_of(env, a[0]);
  if (!h) return nullptr;
  Work *w = new Work;
  w->h = h; w->kind = 2;
  int32_t seed = 0;
  get_i32(env, a[1], "seed", &seed, 0);
  w->law.seed = (uint32_t)seed;
  get_f64(env, a[1], "tone", &w->law.tone, 0.5);
  get_f64(env, a[1], "shoulderLimit", &w->law.shoulder_limit, 0.62 * 3.141592653589793);
  get_f64(env, a[1], "jointLimit", &w->law.joint_limit, 0.42 * 3.141592653589793);
  get_f64(env, a[1], "stiffMin", &w->law.stiff_min, 2000.0);
  get_f64(env, a[1], "stiffMax", &w->law.stiff_max, 400000.0);
  get_f64(env, a[1], "gripMin", &w->law.grip_min, 40.0);
  get_f64(env, a[1], "gripMax", &w->law.grip_max, 900.0);
  get_f64(env, a[1], "slipMax", &w->law.slip_max, 150.0);
  get_f64(env, a[1], "cooldown", &w->law.cooldown, 0.05);
  get_f64(env, a[1], "attachMin", &w->law.attach_min, 0.08);
  get_f64(env, a[1], "releaseBelow", &w->law.release_below, 0.04);
  int32_t cbase = 0, flags = 0;
  get_i32(env, a[1], "creatureBase", &cbase, 0);
  get_i32(env, a[1], "flags", &flags, 0);
  w->law.creature_base = (uint32_t)cbase;
  w->law.flags = (uint32_t)flags;
  napi_get_value_int32(env, a[2], &w->step0);
  napi_get_value_int32(env, a[3], &w->nsteps);
  std::vector<napi_value> keep{a[0]};
  if (argc >= 5) {
    napi_valuetype t;
    napi_typeof(env, a[4], &t);
    if (t != napi_undefined && t != napi_null) {
      int32_t nb, nc, np, ncr;
      kx_counts(h, &nb, &nc, &np, &ncr);
      napi_typedarray_type tt;
      size_t len;
      void *data;
      napi_value ab;
      size_t off;
      bool is = false;
      napi_is_typedarray(env, a[4], &is);
      if (!is || napi_get_typedarray_info(env, a[4], &tt, &len, &data, &ab, &off) != napi_ok ||
          tt != napi_uint32_array || len < (size_t)w->nsteps * (size_t)ncr) {
        delete w;
        napi_throw_type_error(env, nullptr, "kxpbd: events must be a Uint32Array of nsteps*nCreatures");
        return nullptr;
      }
      w->events = static_cast<uint32_t *>(data);
      keep.push_back(a[4]);
    }
  }
  return queue(env, w, keep);
}

// groupPoints(pBody:Int32Array, nBodies, order:Int32Array, range:Int32Array)
napi_value GroupPoints(napi_env env, napi_callback_info info) {
  size_t argc = 4;
  napi_value a[4];
  NAPI_OK(napi_get_cb_info(env, info, &argc, a, nullptr, nullptr));
  if (argc < 4) {
    napi_throw_type_error(env, nullptr, "kxpbd.groupPoints(pBody, nBodies, order, range)");
    return nullptr;
  }
  void *pb, *ord, *rng;
  size_t npb, nord, nrng;
  napi_typedarray_type t1, t2, t3;
  napi_value ab;
  size_t off;
  int32_t nb = 0;
  napi_get_value_int32(env, a[1], &nb);
  if (napi_get_typedarray_info(env, a[0], &t1, &npb, &pb, &ab, &off) != napi_ok ||
      napi_get_typedarray_info(env, a[2], &t2, &nord, &ord, &ab, &off) != napi_ok ||
      napi_get_typedarray_info(env, a[3], &t3, &nrng, &rng, &ab, &off) != napi_ok || t1 != napi_int32_array ||
      t2 != napi_int32_array || t3 != napi_int32_array || nord < npb || nb < 0 || nrng < 2 * (size_t)nb) {
    napi_throw_type_error(env, nullptr, "kxpbd.groupPoints: Int32Arrays of matching sizes expected");
    return nullptr;
  }
  int rc = kx_group_points((int32_t)npb, static_cast<int32_t *>(pb), nb, static_cast<int32_t *>(ord),
                           static_cast<int32_t *>(rng));
  if (rc < 0) return throw_kx(env, rc);
  return nullptr;
}

napi_value LastError(napi_env env, napi_callback_info) {
  napi_value s;
  const char *e = kx_last_error();
  NAPI_OK(napi_create_string_utf8(env, e ? e : "", NAPI_AUTO_LENGTH, &s));
  return s;
}

napi_value AbiVersion(napi_env env, napi_callback_info) {
  napi_value v;
  NAPI_OK(napi_create_int32(env, kx_abi_version(), &v));
  return v;
}

const napi_property_attributes kProp = static_cast<napi_property_attributes>(napi_writable | napi_enumerable);

napi_value Init(napi_env env, napi_value exports) {
  napi_property_descriptor props[] = {
      {"create", nullptr, Create, nullptr, nullptr, nullptr, kProp, nullptr},
      {"destroy", nullptr, Destroy, nullptr, nullptr, nullptr, kProp, nullptr},
      {"counts", nullptr, Counts, nullptr, nullptr, nullptr, kProp, nullptr},
      {"constraintOrder", nullptr, ConstraintOrder, nullptr, nullptr, nullptr, kProp, nullptr},
      {"colorInfo", nullptr, ColorInfo, nullptr, nullptr, nullptr, kProp, nullptr},
      {"upload", nullptr, Upload, nullptr, nullptr, nullptr, kProp, nullptr},
      {"packControls", nullptr, PackControls, nullptr, nullptr, nullptr, kProp, nullptr},
      {"step", nullptr, Step, nullptr, nullptr, nullptr, kProp, nullptr},
      {"substep", nullptr, Substep, nullptr, nullptr, nullptr, kProp, nullptr},
      {"respawn", nullptr, Respawn, nullptr, nullptr, nullptr, kProp, nullptr},
      {"download", nullptr, Download, nullptr, nullptr, nullptr, kProp, nullptr},
      {"rollout", nullptr, Rollout, nullptr, nullptr, nullptr, kProp, nullptr},
      {"readState", nullptr, ReadState, nullptr, nullptr, nullptr, kProp, nullptr},
      {"policyCreate", nullptr, PolicyCreate, nullptr, nullptr, nullptr, kProp, nullptr},
      {"policyWeights", nullptr, PolicyWeights, nullptr, nullptr, nullptr, kProp, nullptr},
      {"policyGenomes", nullptr, PolicyGenomes, nullptr, nullptr, nullptr, kProp, nullptr},
      {"policyForward", nullptr, PolicyForward, nullptr, nullptr, nullptr, kProp, nullptr},
      {"esCreate", nullptr, EsCreate, nullptr, nullptr, nullptr, kProp, nullptr},
      {"esGeneration", nullptr, EsGeneration, nullptr, nullptr, nullptr, kProp, nullptr},
      {"esUpdate", nullptr, EsUpdate, nullptr, nullptr, nullptr, kProp, nullptr},
      {"esGet", nullptr, EsGet, nullptr, nullptr, nullptr, kProp, nullptr},
      {"brainDestroy", nullptr, BrainDestroy, nullptr, nullptr, nullptr, kProp, nullptr},
      {"groupPoints", nullptr, GroupPoints, nullptr, nullptr, nullptr, kProp, nullptr},
      {"lastError", nullptr, LastError, nullptr, nullptr, nullptr, kProp, nullptr},
      {"abiVersion", nullptr, AbiVersion, nullptr, nullptr, nullptr, kProp, nullptr},
  };
  napi_define_properties(env, exports, sizeof(props) / sizeof(props[0]), props);
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
