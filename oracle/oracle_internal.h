/* Internal layout of the CPU oracle — TEST INFRASTRUCTURE ONLY. */
#ifndef KX_ORACLE_INTERNAL_H
#define KX_ORACLE_INTERNAL_H
#include <stdint.h>

/* constraint types (xpbd.js:148-150) */
#define OR_PIN 0
#define OR_ANGLE 1
#define OR_GRIP 2
#define OR_MAX_COLORS 16

/* dead-zones (xpbd.js:77-78) */
#define SLACK_POS 2e-3
#define SLACK_ANG 2e-4

/* BODY constants (kraken.js:22-47) */
#define B_SEG_LEN 13.5
#define B_SHOULDER_R 24.0
#define B_MANTLE_R 18.0
#define B_MANTLE_MASS 20.0
#define B_MANTLE_IF 2.5
#define B_SEG_MASS_PROX 3.2
#define B_SEG_MASS_DIST 1.8
#define B_SEG_THICK_PROX 11.9
#define B_SEG_THICK_DIST 5.9
#define B_JOINT_LIMIT (M_PI * 0.42)
#define B_SHOULDER_LIMIT (M_PI * 0.62)
#define B_SPAWN_CURL 0.28
#define B_MUSCLE_MIN 2000.0
#define B_MUSCLE_MAX 400000.0
#define B_GRIP_MIN 40.0
#define B_GRIP_MAX 900.0
/* sim DEFAULTS (core.js:30-31) */
#define B_SLIP_MAX 150.0
#define B_COOLDOWN 0.05

#define OR_MAX_ARMS 8
#define OR_MAX_SEGS 16
/* control-law PRNG slots: motor m of arm a -> a*16+m, adhesion of arm a -> 256+a */
#define OR_SLOTS_PER_ARM 16
#define OR_SLOT_ADH 256

typedef struct {
  int type, a, b, owner, color, broken, id;
  double lax, lay, lbx, lby, compliance, rest, target, minA, maxA;
  double lambda, lambdaX, lambdaY, breakForce, force;
} OrCons;

typedef struct { int body; double lx, ly, r; } OrPt;

typedef struct {
  int tip, grip, nmotors;
  int motor[OR_MAX_SEGS + 1];   /* creation ids */
  double cd, adhesion;
} OrArm;

typedef struct {
  int mantle, arms, segs, ring, first_body, nbodies;
  OrArm arm[OR_MAX_ARMS];
} OrCreature;

typedef struct OrF32 OrF32;

enum { A_PX, A_PY, A_PPX, A_PPY, A_VX, A_VY, A_ANG, A_ANGP, A_AV, A_IM, A_II, A_FX, A_FY, A_TQ, OR_NARR };

struct OrWorld {
  int S, I;
  double h, linDrag, angDrag, maxSpeed, maxAngSpeed;
  int n, cap;
  double *arr[OR_NARR];
  int nc, ccap;
  OrCons *cons;        /* list order; colour order once coloured */
  int *pos_of_id;      /* creation id -> list position */
  int np, pcap;
  OrPt *pts;
  int nr, has_sdf;
  double *rects;
  double map_w, map_h;
  int colored;
  uint32_t colorOff[OR_MAX_COLORS], colorCnt[OR_MAX_COLORS];
  int ncr, crcap;
  OrCreature *cr;
  double stats_maxSpeed, stats_maxAngSpeed, stats_nan, stats_broken;
  long stats_collisions;
  int step_counter;
  int nev, evcap;
  int *ev;
  OrF32 *f32;
};

#define PX (w->arr[A_PX])
#define PY (w->arr[A_PY])
#define PPX (w->arr[A_PPX])
#define PPY (w->arr[A_PPY])
#define VX (w->arr[A_VX])
#define VY (w->arr[A_VY])
#define ANG (w->arr[A_ANG])
#define ANGP (w->arr[A_ANGP])
#define AV (w->arr[A_AV])
#define IM (w->arr[A_IM])
#define II (w->arr[A_II])
#define FX (w->arr[A_FX])
#define FY (w->arr[A_FY])
#define TQ (w->arr[A_TQ])

double or_js_min(double a, double b);
double or_js_max(double a, double b);
double or_wrap_angle(double a);
double or_angle_delta(double a, double b);
void or_f32_substep(OrWorld *w);
void or_f32_step(OrWorld *w);
void or_log_break(OrWorld *w, const OrCons *c);
OrWorld *or_world_clone(const OrWorld *p);

#endif
