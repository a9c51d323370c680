/*
 * kxpbd — MI355X (gfx950) batched 2-D XPBD step for populations of
 * cephalopod bodies. C ABI of libkxpbd.so.
 *
 * Plain pointers and sizes only; every array is caller-owned host memory in
 * structure-of-arrays form. Functions return 0 on success or a negative
 * KX_E* code; kx_last_error() gives a thread-local message. Calls on one
 * handle must be serialised by the caller; different handles may be used
 * from different threads. Reference citations are relative to the
 * reference repository (lamachen/Kraken_3_GPU).
 *
 * Constraint order: kx_create colours the constraints it is given (those with
 * colour < 0) with the greedy smallest-free-colour rule and then sorts them
 * stably by colour, exactly as ensureColored does (src/physics/xpbd.js:
 * 108-126). Every per-constraint array that later crosses this boundary
 * (upload, pack_controls, download) is in that colour-sorted "handle order";
 * kx_constraint_order() returns the permutation.
 */
#ifndef KXPBD_H
#define KXPBD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KXPBD_ABI_VERSION 5

/* error codes */
#define KX_OK 0
#define KX_EINVAL -22      /* bad argument / failed sanity guard */
#define KX_ENOMEM -12      /* host or device allocation failed */
#define KX_EDEVICE -5      /* HIP runtime error */
#define KX_ENOTSUP -95     /* entry point not available for this handle */

/* constraint types (src/physics/xpbd.js:148-150) */
#define KX_PIN 0
#define KX_ANGLE 1
#define KX_GRIP 2

/* solver paths */
#define KX_PATH_PLAIN 0        /* one launch per colour pass, WGSL-shaped */
#define KX_PATH_PERSISTENT 1   /* one launch per step chunk, creature state in LDS */

/* World description: the static part of an XpbdWorld plus the map.
 * Replaces what gpuworld.create() reads from `cpu` and `map`
 * (src/gpu/gpuworld.js:34-171, writeStatic 209-247). */
typedef struct {
  int32_t n_bodies, n_constraints, n_points, n_rects;
  const float *inv_mass, *inv_inertia;                 /* [n_bodies] */
  const int32_t *c_type, *c_a, *c_b, *c_owner;         /* [n_constraints]; c_b = -1: world anchor */
  const float *c_lax, *c_lay, *c_lbx, *c_lby;          /* anchors (lbx,lby world coords if b<0) */
  const float *c_rest, *c_min, *c_max;                 /* angle rest / limits (+-inf allowed) */
  const int32_t *c_color;                              /* optional; NULL or <0 entries: colour here */
  const int32_t *p_body;                               /* [n_points] */
  const float *p_lx, *p_ly, *p_r;
  const float *rects;                                  /* [n_rects*4] x, y, w, h */
} kx_world_desc;

/* Optional creature structure for the device-side control law
 * (src/sim/core.js:183-220 applyControl). Indices are desc indices. */
typedef struct {
  int32_t n_creatures, n_arms;
  const int32_t *creature_arm_off;   /* [n_creatures+1] */
  const int32_t *arm_tip;            /* [n_arms] tip body */
  const int32_t *arm_grip;           /* [n_arms] grip constraint */
  const int32_t *arm_motor_off;      /* [n_arms+1] */
  const int32_t *motors;             /* motor constraints, shoulder first, then along the chain */
} kx_creature_desc;

/* Solver parameters (src/physics/xpbd.js:46-54; gpuworld.js Params 212-223). */
typedef struct {
  int32_t substeps, iterations;
  double h;                          /* seconds per substep */
  double lin_drag, ang_drag, max_speed, max_ang_speed;
  int32_t path;                      /* KX_PATH_* */
  int32_t lanes_per_creature;        /* persistent path: 8, 16, 32 or 64; 0 = default (16) */
  int32_t waves_per_block;           /* persistent path: 1, 2 or 4; 0 = default (1) */
} kx_params;

/* Body state in SoA (gpuworld.js:250-263 upload; 417-441 mirrorInto).
 * ppx/ppy/angP and fx/fy/tq may be NULL (upload: taken as px/py/ang and 0). */
typedef struct {
  float *px, *py, *vx, *vy, *ang, *av;
  float *ppx, *ppy, *angP;
  float *fx, *fy, *tq;
} kx_body_state;

/* Constraint state in handle order (gpuworld.js:264-281 upload; packControls
 * 331-341; mirrorInto 443-451). Any pointer may be NULL (skipped). */
typedef struct {
  float *lbx, *lby, *compliance, *target, *break_force;
  uint8_t *broken;
  float *lambda, *lambda_x, *lambda_y;
} kx_cons_state;

/* Diagnostics of a download (gpuworld.js:452-455). */
typedef struct {
  double max_speed, max_ang_speed;
  int64_t nan;
} kx_stats;

/* Device-side control law for rollouts: per step s and creature c,
 * u = mulberry32-chain(seed, c, s, slot); motor m of arm a gets target
 * (2u-1)*limit (shoulder_limit for m = 0, joint_limit otherwise), every motor
 * gets compliance 1/(stiff_min + tone*(stiff_max-stiff_min)), and the
 * adhesion a = u(slot 256+a) drives the reflex of core.js:199-218 with
 * break force grip_min + (grip_max-grip_min)*a. */
typedef struct {
  uint32_t seed;
  double tone, shoulder_limit, joint_limit;
  double stiff_min, stiff_max, grip_min, grip_max;
  double slip_max, cooldown, attach_min, release_below;
  uint32_t creature_base;            /* creature c of this handle draws as creature creature_base + c */
  uint32_t flags;                    /* KX_LAW_* */
} kx_law;
/* each tip's break force is drawn once (per creature and arm) instead of
 * following that step's adhesion */
#define KX_LAW_FIXED_GRIP 1u
/* controls come from the handle's control tape (kx_upload_tape) instead of
 * the seeded draws: row r drives the r-th step of a rollout, and applies
 * applyControl (src/sim/core.js:183-220) to its record. seed and tone are
 * unused; limits, stiffness range and the reflex constants still apply. */
#define KX_LAW_TAPE 2u
/* one control-tape record per creature and step, the control struct that
 * decode (src/brain/net.js:114-124) fills: [0] tone, then for arm a < 8
 * [1+3a] shoulder, [2+3a] curl, [3+3a] adhesion */
#define KX_TAPE_WORDS 25

typedef struct kx_handle kx_handle;

/* ── lifecycle (gpuworld.js create 34-171 / destroy 466-472) ─────────── */
int  kx_create(const kx_world_desc *desc, const kx_creature_desc *creatures,
               const kx_params *params, kx_handle **out);
void kx_destroy(kx_handle *h);             /* idempotent on NULL */
const char *kx_last_error(void);
int  kx_abi_version(void);

/* ── layout queries ──────────────────────────────────────────────────── */
int kx_counts(const kx_handle *h, int32_t *n_bodies, int32_t *n_constraints,
              int32_t *n_points, int32_t *n_creatures);
/* order[k] = desc index of the k-th constraint in handle order; color[k] its colour */
int kx_constraint_order(const kx_handle *h, int32_t *order, int32_t *color);
/* colour offsets/counts in handle order (colorInfo, gpuworld.js:286-297) */
int kx_color_info(const kx_handle *h, uint32_t *off16, uint32_t *cnt16);
/* pure host planning of the persistent path's constraint rounds for one
 * creature (local body indices, constraints in colour order) at G lanes:
 * rounds all pin/grip (rtype 0) or all angle (rtype 1), no two constraints
 * sharing a body, every pair that shares a body kept in colour order; an
 * angle round holds at most G constraints, a pin round G / 2 (a lane pair
 * each), or G with KX_PLAN_SINGLE_LANE or-ed into G (one lane each, the form
 * the fused rounds take where lane pairs do not alternate). sched_out[R*G]
 * holds constraint indices or -1; returns R (rounds), or a negative code if
 * max_rounds is too small. */
#define KX_PLAN_SINGLE_LANE 0x100
int kx_plan_rounds(int32_t n_bodies, int32_t n_constraints, const int32_t *c_a, const int32_t *c_b,
                   const int32_t *c_type, int32_t G, int32_t *sched_out, int32_t *rtype_out, int32_t max_rounds);
/* pure host colouring, what kx_create applies (xpbd.js:98-126): colours the
 * constraints whose color_in entry is < 0 (color_in may be NULL) and returns
 * color_out[n_constraints] plus the stable colour sort order_out[k] = desc
 * index of the k-th constraint in handle order. No device needed. */
int kx_color_constraints(int32_t n_bodies, int32_t n_constraints, const int32_t *c_a, const int32_t *c_b,
                         const int32_t *color_in, int32_t *color_out, int32_t *order_out);
/* stable grouping of collision points by body (gpuworld.js:189-207):
 * order[n_points], range[2*n_bodies] = offsets then counts */
int kx_group_points(int32_t n_points, const int32_t *p_body, int32_t n_bodies,
                    int32_t *order, int32_t *range);

/* ── state transfer (gpuworld.js upload 250-298, uploadControls 312-315,
 *    download 385-396) ──────────────────────────────────────────────── */
int kx_upload(kx_handle *h, const kx_body_state *bodies, const kx_cons_state *cons);
int kx_pack_controls(kx_handle *h, const kx_cons_state *controls);
int kx_download(kx_handle *h, kx_body_state *bodies, kx_cons_state *cons, kx_stats *stats);
/* reflex cooldown per creature arm, [n_creatures * 8] doubles */
int kx_get_cooldown(kx_handle *h, double *cd);
int kx_set_cooldown(kx_handle *h, const double *cd);

/* ── stepping (gpuworld.js step 374-382): S substeps + grip break per step;
 *    asynchronous on the handle's stream. ─────────────────────────────── */
int kx_step(kx_handle *h, int32_t nsteps);
/* n substeps without the grip break (gpuworld.js:344-372 substep: reset
 * lambda, integrate, I x (colours, collisions), finalize), on either path's
 * state; asynchronous on the handle's stream. */
int kx_substep(kx_handle *h, int32_t n);
/* nsteps of {control law + reflex, S substeps, grip break}; steps are keyed
 * step0 .. step0+nsteps-1 in the control PRNG. Needs a creature desc.
 * events (optional, device-resident until kx_sync): per step and creature a
 * u32 with bit (3*arm + kind) set for kind 0 attach, 1 release, 2 break. */
int kx_rollout(kx_handle *h, const kx_law *law, int32_t step0, int32_t nsteps,
               uint32_t *events_host);
int kx_sync(kx_handle *h);
/* kx_rollout that allocates nothing. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under graph capture,
 * tight serving loops, harnesses that forbid allocation). The per-step event
 * words go to a static device array of the library (4194304 words = 16 MiB,
 * one per GPU, held until the process ends) instead of the handle's
 * grow-on-demand events buffer; nsteps x creatures beyond that is KX_EINVAL.
 * The words a call uses are cleared first; calls that overlap on one GPU
 * race on the array. Results equal kx_rollout's. */
int kx_rollout_noalloc(kx_handle *h, const kx_law *law, int32_t step0, int32_t nsteps,
                       uint32_t *events_host);
/* Control tape for rollouts under KX_LAW_TAPE: n_rows rows of n_creatures
 * records of KX_TAPE_WORDS floats (row-major). Ordered after the handle's
 * earlier work and copied before return, so the caller may reuse `tape`.
 * A tape rollout of nsteps needs nsteps <= n_rows. Replaces the host side of
 * the reference's per-step control upload (gputrain.js applyControl +
 * gpuworld.js uploadControls 312-315) for K steps at once. */
int kx_upload_tape(kx_handle *h, const float *tape, int32_t n_rows);
/* capture `nsteps` rollout steps into a hipGraph once and replay it `reps`
 * times (step keys advance by nsteps per replay) */
int kx_rollout_graph(kx_handle *h, const kx_law *law, int32_t step0, int32_t nsteps, int32_t reps);
/* average device time (ms) of the last kx_step/kx_rollout launch sequence,
 * measured with HIP events on the handle's stream */
int kx_last_timing(const kx_handle *h, double *ms_total, int32_t *launches);
/* switch path; for the persistent path also its shape (0 keeps the current
 * value): G lanes per creature (64/G creatures per wavefront) and wavefronts
 * per workgroup; rebuilds the per-morphology round schedules for G */
int kx_set_path(kx_handle *h, int32_t path, int32_t lanes_per_creature, int32_t waves_per_block);
/* persistent path: rounds per iteration of each morphology template for the
 * current shape (out[n_templates]); returns the template count */
int kx_schedule_info(const kx_handle *h, int32_t *rounds, int32_t max);
/* persistent launch shape for the current (lanes, waves): info[0] LDS bytes
 * per workgroup, [1] workgroups, [2] resident workgroups per CU, [3]
 * workgroups per CU the launch needs, [4] LDS words of the staged map (0:
 * map read from HBM), [5] cell-bitmap words, [6] template words, [7]
 * per-creature LDS words, [8] 1 if lane counts differ between workgroups,
 * [9] the collision fold in use (KX_FOLD_PARALLEL or KX_FOLD_OWNER), [10]
 * rounds per iteration held in registers (0: the round slots are read from
 * LDS each round; see kx_set_register_slots), [11] rounds per iteration of
 * the fixed schedule run with its slots read from LDS (0: not that kernel),
 * [12] creatures whose template the run-time kernels solve in fused
 * pin/motor rounds (templates whose rounds alternate pins and motors, each
 * motor on the lane of the pin before it with the same two bodies: pins on
 * lane pairs at 16 or 32 lanes, or on single lanes at 8, 16 or 32 where
 * pairs do not alternate; 0 with kx_set_register_slots mode 0 or under
 * [10]/[11]).
 * Returns the number of fields (13). No counterpart in the reference (its
 * path has no persistent kernel). */
int kx_persist_info(const kx_handle *h, int32_t *info, int32_t max);
/* Persistent path: the kernels specialised to the schedule of the default
 * creature (6 arms x 7 segments: 12 rounds at 16 lanes per creature, pins
 * and motors alternating, a joint's hinge and muscle solved in one fused
 * round). enable = 1 (the default): register-held round slots when the
 * launch places at most one wavefront per SIMD (each lane keeps its slot of
 * every round in registers and a round reads only its two poses from LDS),
 * else the same fused rounds with the slots read from LDS; 2: the LDS-read
 * form even where the register form would fit; 0: neither (the kernel that
 * schedules any morphology at run time, every round solved on its own).
 * Templates that do not have that schedule take the run-time kernel, which
 * in modes 1 and 2 solves a template's hinge and muscle rounds fused too
 * when its schedule allows (kx_persist_info field 12). Results are identical
 * bit for bit in every mode. No counterpart in the reference (a launch-shape
 * knob, like waves_per_block). */
int kx_set_register_slots(kx_handle *h, int32_t enable);
/* Persistent path: how a step folds the collision points of the bodies near
 * a rect (xpbd.js:441-470 either way, with identical results).
 * KX_FOLD_PARALLEL tests the near bodies' points in parallel over the
 * creature's lanes and re-runs in order only a body with a correcting point
 * (a candidate list in LDS); KX_FOLD_OWNER folds each near body in order on
 * its owning lane. KX_FOLD_AUTO (the default) takes PARALLEL on an open map
 * (under half of the area within reach of a rect) and OWNER on a dense one
 * (corridors), where most bodies touch a rect most of the time. No
 * counterpart in the reference (a launch-shape knob, like waves_per_block). */
/* State readout for a host policy, in place of mirroring every body (the
 * reference's GPU path reads the whole world back; README.md:693-696 names
 * this compact per-creature stateOut as its next lever). For creature c, at
 * out + c*stride:
 *   [0, 6A+6)   readProprio (src/sim/core.js:129-175) of its A arms: shoulder
 *               angles, curls, tip x/y in the mantle frame, grip loads,
 *               anchored flags, then acceleration along/across, anchored
 *               share, mean extension, speed, angular velocity;
 *   then KX_STATE_EXTRA floats: mantle x, y, angle, vx, vy, av, speed, mean
 *               extension, mean curl, anchored arms, ax, ay (world frame,
 *               unclamped).
 * Computed in binary64 from the device's binary32 state (after the last
 * step or rollout on the handle's stream), returned as float. Each call moves
 * the previous mantle velocity the acceleration inputs use, as readProprio
 * does once per step (zero at kx_create). Needs the creature structure
 * (kx_creature_desc) with each arm's motors running mantle -> segments ->
 * tip. Returns 6*maxA + 6 + KX_STATE_EXTRA (54 for six arms); stride must
 * be at least that. No counterpart in the reference's GPU path. */
#define KX_STATE_EXTRA 12
int kx_read_state(kx_handle *h, float *out, int32_t stride);
/* respawnCreature (src/sim/train.js:188-229) for n creatures in place on the
 * device, no rebuild or full upload: creature creatures[i] gets its mantle at
 * pose[3i], pose[3i+1] with heading pose[3i+2], every arm re-chained from
 * 24 px out along heading + 2*pi*a/A in 13.5 px segments turning 0.28 rad
 * each (float64, rounded to the state's binary32), all velocities 0, every
 * motor target 0, every grip released (broken, lambdaX/Y 0, so its force
 * reads 0), and the readout's previous mantle velocity cleared (readProprio's
 * prevVx/prevVy). Compliances, break limits and reflex cooldowns are kept, as
 * the reference keeps them. Ordered on the handle's stream after earlier
 * work; the host arrays may be reused on return. Ids must be distinct and in
 * range; needs the creature structure with motors running mantle -> segments
 * -> tip (KX_ENOTSUP otherwise). Replaces the host respawn + full upload of
 * a generation start (gputrain.js:130-133 beginGenerationGpu). */
int kx_respawn(kx_handle *h, int32_t n, const int32_t *creatures, const double *pose);
#define KX_FOLD_AUTO 0
#define KX_FOLD_PARALLEL 1
#define KX_FOLD_OWNER 2
int kx_set_collision_fold(kx_handle *h, int32_t fold);
/* Morphology templates (creatures with identical structure share one):
 * t < 0 returns the template count; else out6 = {bodies, constraints, points,
 * creatures, rounds per iteration, lanes per creature} of template t. */
int kx_template_info(const kx_handle *h, int32_t t, int32_t *out6);
/* Template index of every creature (creature order of the description). */
int kx_creature_templates(const kx_handle *h, int32_t *tmpl_of_creature);
/* Lanes per creature per template (0 = the handle's lanes_per_creature), so
 * one persistent launch runs every morphology bucket at its own shape; results
 * do not depend on it. n must equal the template count. */
int kx_set_template_lanes(kx_handle *h, int32_t n, const int32_t *lanes);
/* plain path only: 0 = linear scan over every rect per query (the shape of
 * src/gpu/solver.wgsl:114-147), 1 = the exact cull grid the persistent
 * path always uses */
int kx_set_plain_sdf(kx_handle *h, int32_t use_grid);

/* ── world builder for synthetic workloads (src/body/kraken.js spawnKraken
 *    64-148; src/world/map.js makeMap 22-71, findSpawn 120-128) ────────── */
typedef struct kx_builder kx_builder;
kx_builder *kx_builder_new(void);
void kx_builder_free(kx_builder *b);
int  kx_builder_map(kx_builder *b, double width, double height, int32_t cols, int32_t rows,
                    double density, double wall, uint32_t seed);
int  kx_builder_find_spawn(kx_builder *b, double clearance, uint32_t *rng_state, double *xy);
/* crnSpawns (src/sim/train.js:103-122): pop training spawns drawn from the
 * trainer's stream rng_state (advanced in place; train.js:65-72 seeds it
 * with seed*2654435761): per slot a point in the map's 60 px inner margin
 * whose wall distance lies in [26, ARM_REACH] (500 tries, then an unchecked
 * point) and a heading; with pair_shared, slot 2p+1 repeats slot 2p without
 * drawing. No map built: a 240 x 240 area, untested. out[3*pop] = x, y,
 * angle. Returns the stream calls made, or a negative code. */
int  kx_builder_crn_spawns(kx_builder *b, int32_t pop, int32_t pair_shared, uint32_t *rng_state, double *out);
int  kx_builder_spawn(kx_builder *b, double x, double y, double angle, int32_t arms,
                      int32_t segs, int32_t ring, double curl);
/* views valid until the next builder call */
int  kx_builder_desc(kx_builder *b, kx_world_desc *desc, kx_creature_desc *cr);
/* initial state: px,py,ang of every body (f64 as built; the caller rounds) */
int  kx_builder_state(const kx_builder *b, double *px, double *py, double *ang);
/* initial constraint dynamics in desc order: lbx, lby, compliance, target,
 * break force, broken (spawnKraken: grips broken with anchor at spawn) */
int  kx_builder_cons_init(const kx_builder *b, double *lbx, double *lby, double *compliance,
                          double *target, double *break_force, uint8_t *broken);

/* ── serving: continuous batching of ES rollout requests (the request loop
 *    around the generation loop of src/sim/gputrain.js:72-160, with
 *    the episode of every population member run to its end on the device).
 *    No counterpart in the reference as an API: it runs one population at a
 *    time from one host loop. ─────────────────────────────────────────── */

/* One ES evaluation request: 2*pairs creatures of one morphology
 * (spawnKraken with arms x segs and a mantle ring of `ring` points) at one
 * spawn pose on the server's map, run for `steps` steps. Creature 2p is
 * driven by the policy theta + sigma*eps_p, creature 2p+1 by theta -
 * sigma*eps_p (mirrored sampling as genomeAt does, src/brain/es.js:60-80); theta is
 * drawn from policy_seed and eps_p from (noise_seed, p). The policy is a
 * 64 -> 32 -> 48 tanh MLP of open-loop phase features of the step time,
 * decoded as src/brain/net.js:114-124 decode does (out[0] tone, out[1+3a],
 * out[2+3a], out[3+3a] shoulder, curl, adhesion of arm a); kx_policy_tape
 * gives its records. Outputs in caller memory, written before the
 * request's completion is reported: fitness[2*pairs] (x displacement of each
 * creature's first body over the episode; may be NULL) and final_state
 * (px, py, ang, vx, vy, av blocks of n_bodies floats each; may be NULL or
 * hold final_cap floats, at least 6*n_bodies). */
typedef struct {
  int32_t arms, segs, ring;
  int32_t pairs, steps;
  double x, y, angle;
  uint32_t policy_seed, noise_seed;
  double sigma;
  int64_t tag;
  float *fitness;
  float *final_state;
  int64_t final_cap;
} kx_request;

typedef struct {
  int64_t tag;
  int32_t status;                    /* 0, or a KX_E* code if the request failed */
  int32_t n_creatures, n_bodies;
  double t_submit, t_admit, t_done;  /* seconds on the server clock: arrival, first launch, result ready */
  int64_t device_bytes;              /* device memory the request held (handle arrays and tapes) */
} kx_request_done;

typedef struct {
  int32_t chunk_steps;               /* steps per batched launch (0: 50) */
  int32_t max_active;                /* requests resident on the device at once (0: 64) */
  int32_t host_threads;              /* policy workers (0: hardware threads, at most 16) */
  int64_t hbm_reserve;               /* bytes of free HBM admission leaves untouched */
  kx_params params;                  /* solver parameters (path and shape: persistent, 16 lanes, 1 wave) */
  kx_law law;                        /* limits, stiffness range, reflex constants; KX_LAW_TAPE is implied */
  double map_width, map_height, map_density, map_wall;  /* makeMap (src/world/map.js:22-71) */
  int32_t map_cols, map_rows;
  uint32_t map_seed;
} kx_server_opts;

typedef struct {
  int32_t queued, active, done;      /* waiting for admission, on the device, completed */
  int64_t launches, creature_steps;  /* batched launches so far, creature-steps they ran */
  double now, busy;                  /* server clock; device seconds inside batched launches */
  int64_t peak_active, admit_waits;  /* most requests on the device at once; admissions held back by HBM */
} kx_server_info;

typedef struct kx_server kx_server;
/* starts the server's admission and scheduling threads on the current device */
int  kx_server_create(const kx_server_opts *opts, kx_server **out);
/* queues a request (thread-safe, returns at once); admission is FIFO */
int  kx_server_submit(kx_server *s, const kx_request *req);
/* up to cap completions, waiting at most wait_s seconds for the first */
int  kx_server_poll(kx_server *s, kx_request_done *done, int32_t cap, int32_t *n, double wait_s);
int  kx_server_info_get(kx_server *s, kx_server_info *info);
/* finishes the requests on the device, drops the queued ones, frees all */
void kx_server_destroy(kx_server *s);
/* device bytes admission reserves for a request (state, tables, tapes) */
int64_t kx_request_bytes(const kx_request *req, int32_t chunk_steps);
/* free and total memory of the current device (hipMemGetInfo) */
int  kx_mem_info(int64_t *free_bytes, int64_t *total_bytes);
/* host only: the control-tape records the server feeds request `req` for
 * steps step0 .. step0+nsteps-1, tape[nsteps][2*pairs][KX_TAPE_WORDS] */
int  kx_policy_tape(const kx_request *req, int32_t step0, int32_t nsteps, double step_seconds, float *tape);

/* ── policy network on the device (src/brain/net.js) ─────────────────────
 * A population of n creatures, each with its own genome of
 * kx_policy_weights(n_in, hidden, n_out) floats in randomGenome's layout
 * (net.js:30-50; gputrain.js:55-58 holds them as one flat array). The forward
 * pass is forwardInto (net.js:71-98) for every creature at once: sums in
 * float64 in the reference's order, activations and outputs stored as
 * float32, the layer-2 bias read where net.js:84-86 reads it. hidden and
 * n_out at most 64; the genome, the input row and the activations must fit
 * 64 KiB. */
typedef struct kx_policy kx_policy;
int32_t kx_policy_weights(int32_t n_in, int32_t hidden, int32_t n_out);   /* weightCount, net.js:30-32 */
int  kx_policy_create(int32_t n, int32_t n_in, int32_t hidden, int32_t n_out, kx_policy **out);
int  kx_policy_destroy(kx_policy *p);
/* genomes[n][weights] host <-> device (the population's genomes stay resident) */
int  kx_policy_set_genomes(kx_policy *p, const float *genomes);
int  kx_policy_get_genomes(const kx_policy *p, float *genomes);
/* gputrain.js:77-86 for every creature: x[n][n_in] in (the caller places the
 * previous outputs in the last n_out inputs, gputrain.js:82), out[n][n_out]
 * out; scratch[n][2*hidden] (h1 | h2, may be NULL); controls[n][KX_TAPE_WORDS]
 * (may be NULL): decode (net.js:114-124) of `arms` arms (<= 8) into control-
 * tape records ready for kx_upload_tape; emitEcho stays in out[n_out-1]. */
int  kx_policy_forward(kx_policy *p, const float *x, float *out, float *scratch, float *controls, int32_t arms);
/* One step of gputrain.js:77-100 on the device, with the policy's creatures
 * being the world's (same count, one arm count for all): readProprio of every
 * creature (as kx_read_state), the inputs [proprio | senses | previous
 * outputs] (senses[n][sense_dim] from the host, NULL for zeros;
 * 6*arms + 6 + sense_dim + n_out must equal n_in), the forward pass, decode
 * into control-tape row 0, then a one-step rollout with KX_LAW_TAPE set in a
 * copy of `law` at `step` (applyControl and the step on the device). The
 * outputs stay resident as the next step's recurrent inputs. */
int  kx_policy_drive(kx_policy *p, kx_handle *h, const float *senses, int32_t sense_dim, const kx_law *law,
                     int32_t step);
/* device time of the last forward kernel (HIP events), milliseconds */
int  kx_policy_last_ms(const kx_policy *p, float *ms);
/* reps launches of the forward kernel over the resident inputs and genomes
 * (no host copies); ms = mean kernel time. For measurement. */
int  kx_policy_forward_repeat(kx_policy *p, int32_t reps, float *ms);

/* ── evolution strategy on the device (src/brain/es.js) ────────────────── */
typedef struct {
  double sigma, lr, min_sigma, sigma_decay, beta1, beta2, adam_eps, clip;   /* es.js:23-38 option names */
  int32_t pop_size;                                     /* 1 .. 8192, made even as es.js:37 does */
} kx_es_opts;
typedef struct kx_es kx_es;
/* theta (n_weights floats) may be NULL for zeros (createES, es.js:22-58) */
int  kx_es_create(int32_t n_weights, const kx_es_opts *opts, const float *theta, kx_es **out);
int  kx_es_destroy(kx_es *e);
/* beginGeneration (train.js:154-158): genomeAt (es.js:67-88) for i = 0 ..
 * pop_size-1 into the policy's resident genomes (its n must equal pop_size),
 * drawing from the trainer's mulberry32 stream (train.js:65-72) whose state is
 * *rnd_state; on return *rnd_state is the state after those draws. */
int  kx_es_generation(kx_es *e, kx_policy *p, uint32_t *rnd_state);
/* update (es.js:92-149) from fitness[pop_size] of the last generation;
 * stats4 = {best, mean, sd, sigma} as update returns them (may be NULL) */
int  kx_es_update(kx_es *e, const double *fitness, double *stats4);
/* theta, m, v, best_theta (n_weights floats each, any may be NULL) and
 * scalars3 = {sigma, t, bestEver} (may be NULL) */
int  kx_es_get(const kx_es *e, float *theta, float *m, float *v, float *best_theta, double *scalars3);

#ifdef __cplusplus
}
#endif
#endif
