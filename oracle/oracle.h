/*
 * CPU oracle for the batched 2-D XPBD creature step — TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's CPU algorithm, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * Nothing in the product (kraken_3_gpu_amd/, include/) links or calls it.
 *
 * Reference files followed (paths relative to the reference repository):
 *   src/physics/xpbd.js     f64 world, substep, _solvePin/_solveAngle,
 *                           _solveCollisions, _postGripBreak, colouring
 *   src/body/kraken.js      spawnKraken, tipPosition, attachGrip, releaseGrip,
 *                           toneToCompliance, adhesionToBreakForce
 *   src/world/map.js        mulberry32, makeMap, makeSdf, findSpawn
 *   src/sim/core.js         applyControl (muscle mapping + adhesion reflex)
 *   tools/f32solver.mjs     per-op f32 emulation of src/gpu/solver.wgsl
 *
 * Numerics: compiled with -ffp-contract=off. Math.hypot is restated with
 * V8's normalised Kahan sum (bit-identical to Node). cos/sin/exp come from
 * the host libm, which differs from V8's fdlibm port in the last bit of a few
 * percent of arguments; the golden-trajectory tests bound the effect.
 */
#ifndef KX_ORACLE_H
#define KX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct OrWorld OrWorld;

/* ── PRNG / control law ───────────────────────────────────────────── */
uint32_t or_mb32(uint32_t a);                  /* one mulberry32 output from state a */
void     or_mulberry32_stream(uint32_t seed, int n, double *out);
double   or_ctl_u(uint32_t seed, uint32_t creature, uint32_t step, uint32_t slot);
double   or_js_hypot(double a, double b);

/* ── map ──────────────────────────────────────────────────────────── */
/* makeMap(width,height,{wallThickness,corridorCols,corridorRows,density,seed});
 * writes up to max rects as x,y,w,h; returns the rect count (may exceed max). */
int  or_make_map(double width, double height, int cols, int rows, double density,
                 double wall, uint32_t seed, double *rects, int max);
/* makeSdf fold; returns 1 and d,nx,ny if the rect list is non-empty. */
int  or_sdf(const double *rects, int nr, double x, double y, double *out3);

/* ── world ────────────────────────────────────────────────────────── */
OrWorld *or_world_new(int substeps, int iterations, double h);
void     or_world_free(OrWorld *w);
void     or_world_set_map(OrWorld *w, double width, double height, const double *rects, int nr);
void     or_world_set_limits(OrWorld *w, double linDrag, double angDrag, double maxSpeed, double maxAngSpeed);
/* spawnKraken generalised to arms x segs and a mantle ring of `ring` points.
 * Returns the creature index. */
int      or_spawn_kraken(OrWorld *w, double x, double y, double angle, int arms, int segs, int ring, double curl);
int      or_find_spawn(OrWorld *w, double clearance, uint32_t *rng_state, double *xy);
double   or_rng_next(uint32_t *rng_state);     /* mulberry32 with explicit state */
/* crnSpawns (train.js:103-122) on the world's map (has_map 0: a 240 x 240
 * area, no test); out[3*pop] = x, y, angle; returns the stream calls made */
int      or_crn_spawns(OrWorld *w, int has_map, int pop, int pair_shared, uint32_t *rng_state, double *out);
/* respawnCreature (train.js:188-229): pose chain, velocities, motor targets,
 * grips released; cooldowns untouched */
void     or_respawn(OrWorld *w, int ci, double x, double y, double angle);
void     or_ensure_colored(OrWorld *w);
void     or_permute_cons(OrWorld *w, const int *order);
int      or_counts(const OrWorld *w, int *out4);            /* n, nc, np, ncreatures */
void     or_color_info(const OrWorld *w, uint32_t *off16, uint32_t *cnt16);

/* state access (bodies) */
void or_get_bodies(const OrWorld *w, double *px, double *py, double *ang,
                   double *vx, double *vy, double *av, double *im, double *ii);
void or_set_bodies(OrWorld *w, const double *px, const double *py, const double *ang,
                   const double *vx, const double *vy, const double *av);
void or_get_pred(const OrWorld *w, double *ppx, double *ppy, double *angP);
void or_set_pred(OrWorld *w, const double *ppx, const double *ppy, const double *angP);
/* constraints in current list order:
 * ints  [nc*6]: type, a, b, owner, color, broken
 * dbls  [nc*15]: lax, lay, lbx, lby, compliance, rest, target, minA, maxA,
 *                lambda, lambdaX, lambdaY, breakForce, force, creation_id */
void or_get_cons(const OrWorld *w, int *ints, double *dbls);
void or_set_cons_dynamic(OrWorld *w, const int *broken, const double *lbx, const double *lby,
                         const double *compliance, const double *target, const double *breakForce);
void or_get_points(const OrWorld *w, int *body, double *lxlyr);

/* creature / reflex */
int  or_creature_info(const OrWorld *w, int ci, int *out);  /* mantle, arms, segs, ring, first body, n bodies */
int  or_arm_tip(const OrWorld *w, int ci, int arm);
int  or_arm_grip_pos(const OrWorld *w, int ci, int arm);     /* list position of the grip constraint */
void or_get_cd(const OrWorld *w, double *cd);                /* [ncreatures*8] */
void or_set_cd(OrWorld *w, const double *cd);
void or_attach_grip(OrWorld *w, int ci, int arm, double breakForce);
/* applyControl(c, control) with one shoulder/curl/adhesion per arm */
void or_apply_control(OrWorld *w, int ci, double tone, const double *shoulder,
                      const double *curl, const double *adhesion);
/* the seeded control law: per-motor targets (2u-1)*limit, adhesion u,
 * tone 0.5, then the adhesion reflex; step_index keys the PRNG.
 * Appends reflex events to the event log. */
void or_apply_law(OrWorld *w, uint32_t seed, uint32_t step_index);
void or_apply_law_one(OrWorld *w, uint32_t seed, uint32_t step_index, int creature, uint32_t key);
/* the same law with p5 = {grip_min, grip_max, attach_min, release_below, fixed_grip} */
void or_apply_law_ex(OrWorld *w, uint32_t seed, uint32_t step_index, int creature, uint32_t key, const double *p5);

/* step */
void or_step(OrWorld *w);                /* S substeps + grip break (f64 or f32 emulation) */
void or_substep(OrWorld *w);
void or_solve_collisions(OrWorld *w);    /* one collision pass (f64) */
void or_get_stats(const OrWorld *w, double *out5);  /* maxSpeed, maxAngSpeed, nan, collisions, broken */

/* f32 emulation: snapshot statics, round state; later or_step runs the emulator */
void or_f32ify(OrWorld *w);

/* event log: (step, creature, arm, kind) kind 0 attach, 1 release, 2 break */
int  or_events(const OrWorld *w, int *out, int max);
void or_clear_events(OrWorld *w);
void or_set_step_counter(OrWorld *w, int step);

/* worker-thread throughput baseline: clones `proto` per thread, steps
 * `steps` times under the control law; returns creature-steps per second. */
double or_bench_threads(const OrWorld *proto, int threads, int steps, uint32_t seed);

#ifdef __cplusplus
}
#endif
#endif
