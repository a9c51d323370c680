/*
 * f64 CPU oracle — TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Restates src/physics/xpbd.js, src/body/kraken.js, src/world/map.js and the
 * applyControl reflex of src/sim/core.js in plain C, operation for operation.
 */
#include "oracle.h"
#include "oracle_internal.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ── JS numeric helpers ──────────────────────────────────────────── */

/* Math.hypot for two arguments as V8 computes it: normalise by the larger
 * magnitude and sum the squares with Kahan compensation. */
double or_js_hypot(double a, double b) {
  if (isnan(a) || isnan(b)) {
    if (isinf(a) || isinf(b)) return INFINITY;
    return NAN;
  }
  double x = fabs(a), y = fabs(b);
  double m = x > y ? x : y;
  if (isinf(m)) return INFINITY;
  if (m == 0) return 0;
  double v[2] = {x, y};
  double sum = 0, comp = 0;
  for (int i = 0; i < 2; i++) {
    double n = v[i] / m;
    double summand = n * n - comp;
    double prelim = sum + summand;
    comp = (prelim - sum) - summand;
    sum = prelim;
  }
  return sqrt(sum) * m;
}

double or_js_min(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a == b) return (a == 0 && (signbit(a) || signbit(b))) ? -0.0 : a;
  return a < b ? a : b;
}
double or_js_max(double a, double b) {
  if (a != a || b != b) return NAN;
  if (a == b) return (a == 0 && !(signbit(a) && signbit(b))) ? 0.0 : a;
  return a > b ? a : b;
}
static double clamp01(double v) { return or_js_max(0, or_js_min(1, v)); }

/* xpbd.js:128-134 */
double or_wrap_angle(double a) {
  const double TAU = M_PI * 2;
  double d = fmod(a, TAU);
  if (d > M_PI) d -= TAU;
  else if (d < -M_PI) d += TAU;
  return d;
}
/* xpbd.js:137-143 */
double or_angle_delta(double a, double b) {
  double d = fmod(a - b, M_PI * 2);
  if (!isfinite(d)) return 0;
  if (d > M_PI) d -= M_PI * 2;
  else if (d < -M_PI) d += M_PI * 2;
  return d;
}

/* ── PRNG (map.js:11-20) ─────────────────────────────────────────── */
uint32_t or_mb32(uint32_t a) {
  a += 0x6D2B79F5u;
  uint32_t t = a;
  t = (t ^ (t >> 15)) * (t | 1u);
  t ^= t + (t ^ (t >> 7)) * (t | 61u);
  return t ^ (t >> 14);
}
double or_rng_next(uint32_t *s) {
  *s += 0x6D2B79F5u;
  uint32_t t = *s;
  t = (t ^ (t >> 15)) * (t | 1u);
  t ^= t + (t ^ (t >> 7)) * (t | 61u);
  return (double)(t ^ (t >> 14)) / 4294967296.0;
}
void or_mulberry32_stream(uint32_t seed, int n, double *out) {
  uint32_t s = seed;
  for (int i = 0; i < n; i++) out[i] = or_rng_next(&s);
}
/* Counter-based control PRNG keyed by (seed, creature, step, slot). */
double or_ctl_u(uint32_t seed, uint32_t creature, uint32_t step, uint32_t slot) {
  uint32_t h = or_mb32(seed);
  h = or_mb32(h ^ creature);
  h = or_mb32(h ^ step);
  h = or_mb32(h ^ slot);
  return (double)h / 4294967296.0;
}

/* ── map (map.js:22-71, 91-128) ──────────────────────────────────── */
static int push_rect(double *rects, int n, int max, double x, double y, double w, double h) {
  if (n < max) { rects[4 * n] = x; rects[4 * n + 1] = y; rects[4 * n + 2] = w; rects[4 * n + 3] = h; }
  return n + 1;
}
int or_make_map(double width, double height, int cols, int rows, double density,
                double T, uint32_t seed, double *rects, int max) {
  int n = 0;
  n = push_rect(rects, n, max, 0, 0, width, T);
  n = push_rect(rects, n, max, 0, height - T, width, T);
  n = push_rect(rects, n, max, 0, 0, T, height);
  n = push_rect(rects, n, max, width - T, 0, T, height);
  uint32_t s = seed;
  double cw = width / cols, ch = height / rows;
  for (int col = 1; col < cols; col++) {
    double xl = col * cw;
    for (int row = 0; row < rows; row++) {
      if (or_rng_next(&s) > density) continue;
      double ys = row * ch;
      double gSz = or_js_max(60, or_js_min(ch * (0.28 + or_rng_next(&s) * 0.38), height * 0.4));
      double gStart = ys + or_rng_next(&s) * or_js_max(1, ch - gSz);
      double wt = 10 + floor(or_rng_next(&s) * 10);
      if (gStart > ys + 14) n = push_rect(rects, n, max, xl - wt / 2, ys, wt, gStart - ys);
      double gEnd = gStart + gSz;
      if (gEnd < ys + ch - 14) n = push_rect(rects, n, max, xl - wt / 2, gEnd, wt, ys + ch - gEnd);
    }
  }
  for (int row = 1; row < rows; row++) {
    double yl = row * ch;
    for (int col = 0; col < cols; col++) {
      if (or_rng_next(&s) > density) continue;
      double xs = col * cw;
      double gSz = or_js_max(60, or_js_min(cw * (0.28 + or_rng_next(&s) * 0.38), width * 0.3));
      double gStart = xs + or_rng_next(&s) * or_js_max(1, cw - gSz);
      double wt = 10 + floor(or_rng_next(&s) * 10);
      if (gStart > xs + 14) n = push_rect(rects, n, max, xs, yl - wt / 2, gStart - xs, wt);
      double gEnd = gStart + gSz;
      if (gEnd < xs + cw - 14) n = push_rect(rects, n, max, gEnd, yl - wt / 2, xs + cw - gEnd, wt);
    }
  }
  return n;
}

/* makeSdf fold, including the exterior test against best² (map.js:100-101). */
int or_sdf(const double *R, int nr, double x, double y, double *out) {
  double best = INFINITY, bnx = 0, bny = 0;
  for (int i = 0; i < nr; i++) {
    double rx = R[4 * i], ry = R[4 * i + 1], rw = R[4 * i + 2], rh = R[4 * i + 3];
    double cx = or_js_max(rx, or_js_min(x, rx + rw));
    double cy = or_js_max(ry, or_js_min(y, ry + rh));
    double dx = x - cx, dy = y - cy;
    double d2 = dx * dx + dy * dy;
    if (d2 > 1e-12) {
      if (d2 < best * best) { double d = sqrt(d2); best = d; bnx = dx / d; bny = dy / d; }
    } else {
      double dl = x - rx, dr = rx + rw - x, dt = y - ry, db = ry + rh - y;
      double m = or_js_min(or_js_min(or_js_min(dl, dr), dt), db);
      double d = -m;
      if (d < best) {
        best = d;
        if (m == dl) { bnx = -1; bny = 0; }
        else if (m == dr) { bnx = 1; bny = 0; }
        else if (m == dt) { bnx = 0; bny = -1; }
        else { bnx = 0; bny = 1; }
      }
    }
  }
  if (!isfinite(best)) return 0;
  out[0] = best; out[1] = bnx; out[2] = bny;
  return 1;
}

/* ── world container ─────────────────────────────────────────────── */
OrWorld *or_world_new(int substeps, int iterations, double h) {
  OrWorld *w = (OrWorld *)calloc(1, sizeof(OrWorld));
  w->S = substeps; w->I = iterations; w->h = h;
  w->linDrag = 3.0; w->angDrag = 3.0; w->maxSpeed = 360; w->maxAngSpeed = 18;
  return w;
}
void or_world_free(OrWorld *w) {
  if (!w) return;
  for (int k = 0; k < OR_NARR; k++) free(w->arr[k]);
  free(w->cons); free(w->pts); free(w->rects); free(w->cr); free(w->pos_of_id);
  free(w->ev); free(w->f32);
  free(w);
}
void or_world_set_map(OrWorld *w, double width, double height, const double *rects, int nr) {
  w->map_w = width; w->map_h = height;
  free(w->rects);
  w->rects = (double *)malloc(sizeof(double) * 4 * (nr > 0 ? nr : 1));
  memcpy(w->rects, rects, sizeof(double) * 4 * nr);
  w->nr = nr;
  w->has_sdf = 1;
}
void or_world_set_limits(OrWorld *w, double lin, double ang, double ms, double mas) {
  w->linDrag = lin; w->angDrag = ang; w->maxSpeed = ms; w->maxAngSpeed = mas;
}

static void grow_bodies(OrWorld *w) {
  int cap = w->cap * 2 > 64 ? w->cap * 2 : 64;
  for (int k = 0; k < OR_NARR; k++) {
    double *d = (double *)calloc(cap, sizeof(double));
    if (w->arr[k]) memcpy(d, w->arr[k], sizeof(double) * w->n);
    free(w->arr[k]);
    w->arr[k] = d;
  }
  w->cap = cap;
}
/* xpbd.js:217-228 */
static int add_body(OrWorld *w, double x, double y, double angle, double mass, double inertia) {
  if (w->n >= w->cap) grow_bodies(w);
  int i = w->n++;
  PX[i] = x; PY[i] = y; PPX[i] = x; PPY[i] = y;
  VX[i] = 0; VY[i] = 0; ANG[i] = angle; ANGP[i] = angle; AV[i] = 0;
  IM[i] = mass > 0 ? 1 / mass : 0;
  II[i] = inertia > 0 ? 1 / inertia : 0;
  FX[i] = 0; FY[i] = 0; TQ[i] = 0;
  return i;
}
static void add_point(OrWorld *w, int body, double lx, double ly, double r) {
  if (w->np >= w->pcap) {
    w->pcap = w->pcap ? w->pcap * 2 : 256;
    w->pts = (OrPt *)realloc(w->pts, sizeof(OrPt) * w->pcap);
  }
  OrPt *p = &w->pts[w->np++];
  p->body = body; p->lx = lx; p->ly = ly; p->r = r;
}
/* makeConstraint (xpbd.js:160-179) + addConstraint; returns the creation id */
static int add_cons(OrWorld *w, int type, int a, int b, double lax, double lay,
                    double lbx, double lby, double compliance) {
  if (w->nc >= w->ccap) {
    w->ccap = w->ccap ? w->ccap * 2 : 256;
    w->cons = (OrCons *)realloc(w->cons, sizeof(OrCons) * w->ccap);
    w->pos_of_id = (int *)realloc(w->pos_of_id, sizeof(int) * w->ccap);
  }
  int id = w->nc++;
  OrCons *c = &w->cons[id];
  memset(c, 0, sizeof(*c));
  c->type = type; c->a = a; c->b = b;
  c->lax = lax; c->lay = lay; c->lbx = lbx; c->lby = lby; c->compliance = compliance;
  c->breakForce = INFINITY;
  c->minA = -INFINITY; c->maxA = INFINITY;
  c->owner = -1; c->color = -1; c->id = id;
  w->pos_of_id[id] = id;
  w->colored = 0;
  return id;
}
#define CONS_ID(w, id) (&(w)->cons[(w)->pos_of_id[id]])

/* ── spawnKraken (kraken.js:64-148), generalised morphology ──────── */
int or_spawn_kraken(OrWorld *w, double x, double y, double angle, int arms, int segs, int ring, double curl) {
  if (arms > OR_MAX_ARMS || segs > OR_MAX_SEGS || arms < 1 || segs < 1) return -1;
  if (w->ncr >= w->crcap) {
    w->crcap = w->crcap ? w->crcap * 2 : 16;
    w->cr = (OrCreature *)realloc(w->cr, sizeof(OrCreature) * w->crcap);
  }
  OrCreature *K = &w->cr[w->ncr];
  memset(K, 0, sizeof(*K));
  K->arms = arms; K->segs = segs; K->ring = ring;
  K->first_body = w->n;
  const double mantleI = (B_MANTLE_MASS * B_MANTLE_R * B_MANTLE_R / 2) * B_MANTLE_IF;
  int mantle = add_body(w, x, y, angle, B_MANTLE_MASS, mantleI);
  K->mantle = mantle;
  for (int i = 0; i < ring; i++) {
    double a = ((double)i / ring) * M_PI * 2;
    add_point(w, mantle, cos(a) * B_MANTLE_R * 0.7, sin(a) * B_MANTLE_R * 0.7, B_MANTLE_R * 0.45);
  }
  for (int a = 0; a < arms; a++) {
    double relAng = ((double)a / arms) * M_PI * 2;
    double th0 = angle + relAng;
    int sg[OR_MAX_SEGS];
    double jx = x + cos(th0) * B_SHOULDER_R;
    double jy = y + sin(th0) * B_SHOULDER_R;
    double th = th0;
    for (int i = 0; i < segs; i++) {
      double t = segs > 1 ? (double)i / (segs - 1) : 0;
      double mass = B_SEG_MASS_PROX + (B_SEG_MASS_DIST - B_SEG_MASS_PROX) * t;
      double thick = B_SEG_THICK_PROX + (B_SEG_THICK_DIST - B_SEG_THICK_PROX) * t;
      double ux = cos(th), uy = sin(th);
      double cx = jx + ux * B_SEG_LEN * 0.5;
      double cy = jy + uy * B_SEG_LEN * 0.5;
      double inertia = mass * (B_SEG_LEN * B_SEG_LEN + thick * thick) / 12;
      int b = add_body(w, cx, cy, th, mass, inertia);
      double r = thick * 0.5;
      add_point(w, b, -B_SEG_LEN * 0.5 + r * 0.4, 0, r);
      add_point(w, b, B_SEG_LEN * 0.5 - r * 0.4, 0, r);
      sg[i] = b;
      jx = cx + ux * B_SEG_LEN * 0.5;
      jy = cy + uy * B_SEG_LEN * 0.5;
      th += curl;
    }
    OrArm *arm = &K->arm[a];
    arm->tip = sg[segs - 1];
    arm->nmotors = 0;
    add_cons(w, OR_PIN, mantle, sg[0], cos(relAng) * B_SHOULDER_R, sin(relAng) * B_SHOULDER_R,
             -B_SEG_LEN * 0.5, 0, 0);
    int ms = add_cons(w, OR_ANGLE, mantle, sg[0], 0, 0, 0, 0, 1 / B_MUSCLE_MIN);
    OrCons *m = CONS_ID(w, ms);
    m->rest = relAng; m->target = 0; m->minA = -B_SHOULDER_LIMIT; m->maxA = B_SHOULDER_LIMIT; m->owner = a;
    arm->motor[arm->nmotors++] = ms;
    for (int i = 0; i < segs - 1; i++) {
      add_cons(w, OR_PIN, sg[i], sg[i + 1], B_SEG_LEN * 0.5, 0, -B_SEG_LEN * 0.5, 0, 0);
      int mj = add_cons(w, OR_ANGLE, sg[i], sg[i + 1], 0, 0, 0, 0, 1 / B_MUSCLE_MIN);
      OrCons *mm = CONS_ID(w, mj);
      mm->rest = curl; mm->target = 0; mm->minA = -B_JOINT_LIMIT; mm->maxA = B_JOINT_LIMIT; mm->owner = a;
      arm->motor[arm->nmotors++] = mj;
    }
    int g = add_cons(w, OR_GRIP, arm->tip, -1, B_SEG_LEN * 0.5, 0, x, y, 0);
    OrCons *gc = CONS_ID(w, g);
    gc->broken = 1; gc->breakForce = 0; gc->owner = a;
    arm->grip = g;
    arm->cd = 0; arm->adhesion = 0;
  }
  K->nbodies = w->n - K->first_body;
  return w->ncr++;
}

/* findSpawn (map.js:120-128) with the world's rect list as the map */
int or_find_spawn(OrWorld *w, double clearance, uint32_t *s, double *xy) {
  double hit[3];
  double W = w->map_w, H = w->map_h;
  for (int i = 0; i < 500; i++) {
    double x = clearance + or_rng_next(s) * or_js_max(1, W - clearance * 2);
    double y = clearance + or_rng_next(s) * or_js_max(1, H - clearance * 2);
    if (or_sdf(w->rects, w->nr, x, y, hit) && hit[0] >= clearance) { xy[0] = x; xy[1] = y; return 1; }
  }
  xy[0] = W * 0.5; xy[1] = H * 0.5;
  return 0;
}

/* crnSpawns (train.js:103-122): per slot, up to 500 draws of a point in the
 * map's 60 px inner margin whose fold distance lies in [SPAWN_CLEAR,
 * ARM_REACH] (a wall within reach, the mantle clear), else an unchecked draw;
 * then the heading. With pair_shared the odd slot repeats its even
 * neighbour's spawn and draws nothing. Without a map (has_map 0) the area is
 * 240 x 240 and no point is tested. out: x, y, angle per slot. Returns the
 * number of stream calls made. */
#define OR_SPAWN_CLEAR 26.0
#define OR_ARM_REACH (B_SHOULDER_R + 7 * B_SEG_LEN)
int or_crn_spawns(OrWorld *w, int has_map, int pop, int pair_shared, uint32_t *s, double *out) {
  double W = has_map ? w->map_w : 240, H = has_map ? w->map_h : 240, hit[3];
  int calls = 0;
  for (int i = 0; i < pop; i++) {
    if (pair_shared && (i & 1)) { memcpy(out + 3 * i, out + 3 * (i - 1), 3 * sizeof(double)); continue; }
    int found = 0;
    double x = 0, y = 0;
    if (has_map) {
      for (int t = 0; t < 500 && !found; t++) {
        x = 60 + or_rng_next(s) * or_js_max(1, W - 120);
        y = 60 + or_rng_next(s) * or_js_max(1, H - 120);
        calls += 2;
        if (or_sdf(w->rects, w->nr, x, y, hit) && hit[0] >= OR_SPAWN_CLEAR && hit[0] <= OR_ARM_REACH) found = 1;
      }
    }
    if (!found) {
      x = 60 + or_rng_next(s) * or_js_max(1, W - 120);
      y = 60 + or_rng_next(s) * or_js_max(1, H - 120);
      calls += 2;
    }
    out[3 * i] = x; out[3 * i + 1] = y; out[3 * i + 2] = or_rng_next(s) * M_PI * 2;
    calls++;
  }
  return calls;
}

/* respawnCreature (train.js:188-229), physics part: the mantle at (x, y,
 * angle), every arm re-chained from 24 px out along angle + 2*pi*a/arms in
 * 13.5 px segments turning 0.28 rad each; poses and predictions set,
 * velocities zeroed; every motor target 0; every grip broken with force and
 * lambdaX/Y cleared. Reflex cooldowns, compliances and the other
 * multipliers stay as they were (the reference leaves them). */
void or_respawn(OrWorld *w, int ci, double x, double y, double angle) {
  OrCreature *K = &w->cr[ci];
  int m = K->mantle;
  PX[m] = x; PY[m] = y; PPX[m] = x; PPY[m] = y; VX[m] = 0; VY[m] = 0;
  ANG[m] = angle; ANGP[m] = angle; AV[m] = 0;
  for (int a = 0; a < K->arms; a++) {
    OrArm *arm = &K->arm[a];
    double relAng = ((double)a / K->arms) * M_PI * 2;
    double th0 = angle + relAng;
    double jx = x + cos(th0) * 24, jy = y + sin(th0) * 24;
    double th = th0;
    for (int i = 0; i < K->segs; i++) {
      int b = K->first_body + 1 + a * K->segs + i;
      double ux = cos(th), uy = sin(th);
      double cx = jx + ux * 13.5 * 0.5, cy = jy + uy * 13.5 * 0.5;
      PX[b] = cx; PY[b] = cy; PPX[b] = cx; PPY[b] = cy;
      VX[b] = 0; VY[b] = 0;
      ANG[b] = th; ANGP[b] = th; AV[b] = 0;
      jx = cx + ux * 13.5 * 0.5; jy = cy + uy * 13.5 * 0.5;
      th += 0.28;
    }
    for (int j = 0; j < arm->nmotors; j++) CONS_ID(w, arm->motor[j])->target = 0;
    OrCons *g = CONS_ID(w, arm->grip);
    g->broken = 1; g->force = 0; g->lambdaX = 0; g->lambdaY = 0;
  }
}

/* ── colouring (xpbd.js:98-126) ──────────────────────────────────── */
void or_ensure_colored(OrWorld *w) {
  if (w->colored) return;
  int msz = w->n * 2 > 1024 ? w->n * 2 : 1024;
  uint32_t *mask = (uint32_t *)calloc(msz, sizeof(uint32_t));
  for (int k = 0; k < w->nc; k++) {
    OrCons *c = &w->cons[k];
    if (c->color < 0) {
      uint32_t used = mask[c->a] | (c->b >= 0 ? mask[c->b] : 0);
      int col = 0;
      while (used & (1u << col)) col++;
      mask[c->a] |= 1u << col;
      if (c->b >= 0) mask[c->b] |= 1u << col;
      c->color = col;
    }
  }
  free(mask);
  /* stable bucket sort by colour */
  OrCons *sorted = (OrCons *)malloc(sizeof(OrCons) * (w->nc > 0 ? w->nc : 1));
  int n = 0;
  for (int col = 0; col < OR_MAX_COLORS; col++) {
    w->colorOff[col] = n;
    int cnt = 0;
    for (int k = 0; k < w->nc; k++)
      if ((w->cons[k].color | 0) == col) { sorted[n++] = w->cons[k]; cnt++; }
    w->colorCnt[col] = cnt;
  }
  memcpy(w->cons, sorted, sizeof(OrCons) * w->nc);
  free(sorted);
  for (int k = 0; k < w->nc; k++) w->pos_of_id[w->cons[k].id] = k;
  w->colored = 1;
}

/* reorder the constraint list: new position k holds old position order[k] */
void or_permute_cons(OrWorld *w, const int *order) {
  OrCons *tmp = (OrCons *)malloc(sizeof(OrCons) * (w->nc > 0 ? w->nc : 1));
  for (int k = 0; k < w->nc; k++) tmp[k] = w->cons[order[k]];
  memcpy(w->cons, tmp, sizeof(OrCons) * w->nc);
  free(tmp);
  for (int k = 0; k < w->nc; k++) w->pos_of_id[w->cons[k].id] = k;
}

int or_counts(const OrWorld *w, int *o) { o[0] = w->n; o[1] = w->nc; o[2] = w->np; o[3] = w->ncr; return w->n; }
void or_color_info(const OrWorld *w, uint32_t *off, uint32_t *cnt) {
  for (int c = 0; c < 16; c++) { off[c] = w->colorOff[c]; cnt[c] = w->colorCnt[c]; }
}

/* ── state access ────────────────────────────────────────────────── */
#define CPY(dst, src) do { if (dst) memcpy((dst), (src), sizeof(double) * w->n); } while (0)
void or_get_bodies(const OrWorld *w, double *px, double *py, double *ang, double *vx, double *vy,
                   double *av, double *im, double *ii) {
  CPY(px, PX); CPY(py, PY); CPY(ang, ANG); CPY(vx, VX); CPY(vy, VY); CPY(av, AV); CPY(im, IM); CPY(ii, II);
}
void or_set_bodies(OrWorld *w, const double *px, const double *py, const double *ang,
                   const double *vx, const double *vy, const double *av) {
  CPY(PX, px); CPY(PY, py); CPY(ANG, ang); CPY(VX, vx); CPY(VY, vy); CPY(AV, av);
  CPY(PPX, px); CPY(PPY, py); CPY(ANGP, ang);
}
void or_get_pred(const OrWorld *w, double *ppx, double *ppy, double *angP) {
  CPY(ppx, PPX); CPY(ppy, PPY); CPY(angP, ANGP);
}
void or_set_pred(OrWorld *w, const double *ppx, const double *ppy, const double *angP) {
  CPY(PPX, ppx); CPY(PPY, ppy); CPY(ANGP, angP);
}
void or_get_cons(const OrWorld *w, int *I, double *D) {
  for (int k = 0; k < w->nc; k++) {
    const OrCons *c = &w->cons[k];
    int *i = I + 6 * k; double *d = D + 15 * k;
    i[0] = c->type; i[1] = c->a; i[2] = c->b; i[3] = c->owner; i[4] = c->color; i[5] = c->broken;
    d[0] = c->lax; d[1] = c->lay; d[2] = c->lbx; d[3] = c->lby; d[4] = c->compliance; d[5] = c->rest;
    d[6] = c->target; d[7] = c->minA; d[8] = c->maxA; d[9] = c->lambda; d[10] = c->lambdaX;
    d[11] = c->lambdaY; d[12] = c->breakForce; d[13] = c->force; d[14] = c->id;
  }
}
void or_set_cons_dynamic(OrWorld *w, const int *broken, const double *lbx, const double *lby,
                         const double *comp, const double *target, const double *bf) {
  for (int k = 0; k < w->nc; k++) {
    OrCons *c = &w->cons[k];
    if (broken) c->broken = broken[k];
    if (lbx) c->lbx = lbx[k];
    if (lby) c->lby = lby[k];
    if (comp) c->compliance = comp[k];
    if (target) c->target = target[k];
    if (bf) c->breakForce = bf[k];
  }
}
void or_get_points(const OrWorld *w, int *body, double *d) {
  for (int i = 0; i < w->np; i++) {
    body[i] = w->pts[i].body;
    d[3 * i] = w->pts[i].lx; d[3 * i + 1] = w->pts[i].ly; d[3 * i + 2] = w->pts[i].r;
  }
}
int or_creature_info(const OrWorld *w, int ci, int *o) {
  const OrCreature *K = &w->cr[ci];
  o[0] = K->mantle; o[1] = K->arms; o[2] = K->segs; o[3] = K->ring; o[4] = K->first_body; o[5] = K->nbodies;
  return 0;
}
int or_arm_tip(const OrWorld *w, int ci, int a) { return w->cr[ci].arm[a].tip; }
int or_arm_grip_pos(const OrWorld *w, int ci, int a) { return w->pos_of_id[w->cr[ci].arm[a].grip]; }
void or_get_cd(const OrWorld *w, double *cd) {
  for (int c = 0; c < w->ncr; c++) for (int a = 0; a < OR_MAX_ARMS; a++) cd[c * OR_MAX_ARMS + a] = w->cr[c].arm[a].cd;
}
void or_set_cd(OrWorld *w, const double *cd) {
  for (int c = 0; c < w->ncr; c++) for (int a = 0; a < OR_MAX_ARMS; a++) w->cr[c].arm[a].cd = cd[c * OR_MAX_ARMS + a];
}

/* ── control + reflex (core.js:183-220; kraken.js:150-198) ───────── */
static double tone_to_compliance(double tone) {
  double t = or_js_max(0, or_js_min(1, tone));
  double stiff = B_MUSCLE_MIN + t * (B_MUSCLE_MAX - B_MUSCLE_MIN);
  return 1 / stiff;
}
static double adhesion_to_break_force(double a) {
  double c = or_js_max(0, or_js_min(1, a));
  return B_GRIP_MIN + (B_GRIP_MAX - B_GRIP_MIN) * c;
}
void or_attach_grip(OrWorld *w, int ci, int a, double breakForce) {
  OrArm *arm = &w->cr[ci].arm[a];
  int b = arm->tip;
  double L = B_SEG_LEN * 0.5;
  double c = cos(ANG[b]), s = sin(ANG[b]);
  double tx = PX[b] + L * c, ty = PY[b] + L * s;
  OrCons *g = CONS_ID(w, arm->grip);
  g->lbx = tx; g->lby = ty;
  g->broken = 0; g->breakForce = breakForce;
  g->lambdaX = 0; g->lambdaY = 0; g->force = 0;
}
static void log_event(OrWorld *w, int ci, int a, int kind) {
  if (w->nev >= w->evcap) {
    w->evcap = w->evcap ? w->evcap * 2 : 1024;
    w->ev = (int *)realloc(w->ev, sizeof(int) * 4 * w->evcap);
  }
  int *e = w->ev + 4 * w->nev++;
  e[0] = w->step_counter; e[1] = ci; e[2] = a; e[3] = kind;
}
/* adhesion reflex for one arm (core.js:199-218); p carries the grip range
 * and thresholds (the reference's constants unless a test law overrides
 * them) and grip_u the once-per-tip draw used when p->fixed_grip is set */
typedef struct { double grip_min, grip_max, attach_min, release_below; int fixed_grip; } OrReflexParams;
static const OrReflexParams REF_REFLEX = {B_GRIP_MIN, B_GRIP_MAX, 0.08, 0.04, 0};

static void reflex_p(OrWorld *w, int ci, int a, double adhesion, double grip_u, const OrReflexParams *p) {
  OrArm *arm = &w->cr[ci].arm[a];
  arm->adhesion = clamp01(adhesion);
  if (arm->cd > 0) arm->cd -= w->h * w->S;
  OrCons *g = CONS_ID(w, arm->grip);
  double ga = p->fixed_grip ? clamp01(grip_u) : arm->adhesion;
  double bf = p->grip_min + (p->grip_max - p->grip_min) * ga;
  if (g->broken) {
    if (arm->adhesion > p->attach_min && arm->cd <= 0) {
      double slip = or_js_hypot(VX[arm->tip], VY[arm->tip]);
      if (slip < B_SLIP_MAX) {
        or_attach_grip(w, ci, a, bf);
        log_event(w, ci, a, 0);
      } else {
        arm->cd = B_COOLDOWN;
      }
    }
  } else {
    g->breakForce = bf;
    if (arm->adhesion < p->release_below) { g->broken = 1; g->force = 0; log_event(w, ci, a, 1); }
  }
}
static void reflex(OrWorld *w, int ci, int a, double adhesion) { reflex_p(w, ci, a, adhesion, 0.0, &REF_REFLEX); }

void or_apply_control(OrWorld *w, int ci, double tone, const double *shoulder, const double *curl,
                      const double *adhesion) {
  OrCreature *K = &w->cr[ci];
  double comp = tone_to_compliance(or_js_max(0, or_js_min(1, tone)));
  for (int a = 0; a < K->arms; a++) {
    OrArm *arm = &K->arm[a];
    for (int m = 0; m < arm->nmotors; m++) CONS_ID(w, arm->motor[m])->compliance = comp;
    CONS_ID(w, arm->motor[0])->target = or_js_max(-B_SHOULDER_LIMIT, or_js_min(B_SHOULDER_LIMIT, shoulder[a]));
    for (int j = 1; j < arm->nmotors; j++)
      CONS_ID(w, arm->motor[j])->target = or_js_max(-B_JOINT_LIMIT, or_js_min(B_JOINT_LIMIT, curl[a]));
    reflex(w, ci, a, adhesion[a]);
  }
}
/* the control law for creature ci, keyed in the PRNG as creature `key`
 * (the batch index when ci lives in a sub-world); p = {grip_min, grip_max,
 * attach_min, release_below, fixed_grip} */
static void apply_law_p(OrWorld *w, uint32_t seed, uint32_t step, int ci, uint32_t key, const OrReflexParams *p);
void or_apply_law_ex(OrWorld *w, uint32_t seed, uint32_t step, int ci, uint32_t key, const double *p5) {
  OrReflexParams p = {p5[0], p5[1], p5[2], p5[3], p5[4] != 0.0};
  apply_law_p(w, seed, step, ci, key, &p);
}
void or_apply_law_one(OrWorld *w, uint32_t seed, uint32_t step, int ci, uint32_t key) {
  apply_law_p(w, seed, step, ci, key, &REF_REFLEX);
}
static void apply_law_p(OrWorld *w, uint32_t seed, uint32_t step, int ci, uint32_t key, const OrReflexParams *p) {
  double comp = tone_to_compliance(0.5);
  OrCreature *K = &w->cr[ci];
  for (int a = 0; a < K->arms; a++) {
    OrArm *arm = &K->arm[a];
    for (int m = 0; m < arm->nmotors; m++) {
      double u = or_ctl_u(seed, key, step, a * OR_SLOTS_PER_ARM + m);
      double lim = m == 0 ? B_SHOULDER_LIMIT : B_JOINT_LIMIT;
      double t = (2 * u - 1) * lim;
      OrCons *c = CONS_ID(w, arm->motor[m]);
      c->compliance = comp;
      c->target = or_js_max(-lim, or_js_min(lim, t));
    }
    reflex_p(w, ci, a, or_ctl_u(seed, key, step, OR_SLOT_ADH + a),
             p->fixed_grip ? or_ctl_u(seed, key, 0xFFFFFFFFu, OR_SLOT_ADH + a) : 0.0, p);
  }
}
void or_apply_law(OrWorld *w, uint32_t seed, uint32_t step) {
  for (int ci = 0; ci < w->ncr; ci++) or_apply_law_one(w, seed, step, ci, (uint32_t)ci);
}

/* ── solver (xpbd.js:257-504) ────────────────────────────────────── */
static void solve_pin(OrWorld *w, OrCons *c, int first) {
  if (first) { c->lambdaX = 0; c->lambdaY = 0; }
  int A = c->a, B = c->b;
  double ca = cos(ANGP[A]), sa = sin(ANGP[A]);
  double rax = c->lax * ca - c->lay * sa, ray = c->lax * sa + c->lay * ca;
  double gax = PPX[A] + rax, gay = PPY[A] + ray;
  double gbx, gby, rbx = 0, rby = 0, wB = 0, wIB = 0;
  if (B >= 0) {
    double cb = cos(ANGP[B]), sb = sin(ANGP[B]);
    rbx = c->lbx * cb - c->lby * sb; rby = c->lbx * sb + c->lby * cb;
    gbx = PPX[B] + rbx; gby = PPY[B] + rby;
    wB = IM[B]; wIB = II[B];
  } else {
    gbx = c->lbx; gby = c->lby;
  }
  double Cx = gax - gbx, Cy = gay - gby;
  if (!isfinite(Cx) || !isfinite(Cy)) return;
  if (Cx * Cx + Cy * Cy < SLACK_POS * SLACK_POS) return;
  double pax = -ray, pay = rax;
  double pbx = -rby, pby = rbx;
  double wA = IM[A], wIA = II[A];
  double wSum = wA + wB;
  double k00 = wSum + wIA * pax * pax + wIB * pbx * pbx;
  double k01 = wIA * pax * pay + wIB * pbx * pby;
  double k11 = wSum + wIA * pay * pay + wIB * pby * pby;
  double det = k00 * k11 - k01 * k01;
  if (fabs(det) < 1e-12) return;
  double alphaTilde = c->compliance / (w->h * w->h);
  double bx = Cx + alphaTilde * c->lambdaX;
  double by = Cy + alphaTilde * c->lambdaY;
  double i00 = k11 / det, i11 = k00 / det, i01 = -k01 / det;
  double dlx = -(i00 * bx + i01 * by);
  double dly = -(i01 * bx + i11 * by);
  if (!isfinite(dlx) || !isfinite(dly)) return;
  c->lambdaX += dlx; c->lambdaY += dly;
  PPX[A] += wA * dlx; PPY[A] += wA * dly;
  ANGP[A] += wIA * (pax * dlx + pay * dly);
  if (B >= 0) {
    PPX[B] -= wB * dlx; PPY[B] -= wB * dly;
    ANGP[B] -= wIB * (pbx * dlx + pby * dly);
  }
}

static void solve_angle(OrWorld *w, OrCons *c, int first) {
  if (first) c->lambda = 0;
  int A = c->a, B = c->b;
  double wA = II[A], wB = B >= 0 ? II[B] : 0;
  double wSum = wA + wB;
  if (wSum <= 0) return;
  double rel = B >= 0 ? or_angle_delta(ANGP[B], ANGP[A]) : ANGP[A];
  double dev = or_angle_delta(rel, c->rest);
  double C = or_angle_delta(dev, c->target);
  if (!isfinite(C)) return;
  if (fabs(C) >= SLACK_ANG) {
    double alphaTilde = c->compliance / (w->h * w->h);
    double dl = -(C + alphaTilde * c->lambda) / (wSum + alphaTilde);
    if (!isfinite(dl)) return;
    c->lambda += dl;
    if (B >= 0) { ANGP[B] += wB * dl; ANGP[A] -= wA * dl; }
    else { ANGP[A] += wA * dl; }
  }
  double rel2 = B >= 0 ? or_angle_delta(ANGP[B], ANGP[A]) : ANGP[A];
  double dev2 = or_angle_delta(rel2, c->rest);
  double over = 0;
  if (dev2 > c->maxA) over = dev2 - c->maxA;
  else if (dev2 < c->minA) over = dev2 - c->minA;
  if (fabs(over) >= SLACK_ANG && wSum > 0) {
    double dl2 = -over / wSum;
    if (B >= 0) { ANGP[B] += wB * dl2; ANGP[A] -= wA * dl2; }
    else { ANGP[A] += wA * dl2; }
  }
}

void or_solve_collisions(OrWorld *w) {
  if (!w->has_sdf) return;
  double hit[3];
  for (int k = 0; k < w->np; k++) {
    const OrPt *cp = &w->pts[k];
    int B = cp->body;
    double ca = cos(ANGP[B]), sa = sin(ANGP[B]);
    double wx = PPX[B] + cp->lx * ca - cp->ly * sa;
    double wy = PPY[B] + cp->lx * sa + cp->ly * ca;
    if (!or_sdf(w->rects, w->nr, wx, wy, hit)) continue;
    double pen = cp->r - hit[0];
    if (pen <= SLACK_POS) continue;
    double wm = IM[B];
    if (wm == 0) continue;
    PPX[B] += hit[1] * pen;
    PPY[B] += hit[2] * pen;
    w->stats_collisions++;
  }
}

static void measure(OrWorld *w) {
  double mv = 0, mw = 0; int nan = 0;
  for (int i = 0; i < w->n; i++) {
    double s = or_js_hypot(VX[i], VY[i]);
    if (s > mv) mv = s;
    double a = fabs(AV[i]);
    if (a > mw) mw = a;
    if (!isfinite(PX[i]) || !isfinite(PY[i]) || !isfinite(ANG[i])) nan++;
  }
  w->stats_maxSpeed = mv; w->stats_maxAngSpeed = mw; w->stats_nan = nan;
}

static void substep_f64(OrWorld *w) {
  const double h = w->h;
  const int n = w->n;
  double linDamp = exp(-w->linDrag * h);
  double angDamp = exp(-w->angDrag * h);
  for (int i = 0; i < n; i++) {
    if (IM[i] == 0 && II[i] == 0) {
      PPX[i] = PX[i]; PPY[i] = PY[i]; ANGP[i] = ANG[i];
      VX[i] = 0; VY[i] = 0; AV[i] = 0;
      continue;
    }
    VX[i] = (VX[i] + h * IM[i] * FX[i]) * linDamp;
    VY[i] = (VY[i] + h * IM[i] * FY[i]) * linDamp;
    AV[i] = (AV[i] + h * II[i] * TQ[i]) * angDamp;
    double sp = or_js_hypot(VX[i], VY[i]);
    if (sp > w->maxSpeed) { double k = w->maxSpeed / sp; VX[i] *= k; VY[i] *= k; }
    if (AV[i] > w->maxAngSpeed) AV[i] = w->maxAngSpeed;
    else if (AV[i] < -w->maxAngSpeed) AV[i] = -w->maxAngSpeed;
    if (!isfinite(VX[i]) || !isfinite(VY[i])) { VX[i] = 0; VY[i] = 0; w->stats_nan++; }
    if (!isfinite(AV[i])) { AV[i] = 0; w->stats_nan++; }
    PPX[i] = PX[i] + h * VX[i];
    PPY[i] = PY[i] + h * VY[i];
    ANGP[i] = ANG[i] + h * AV[i];
    FX[i] = 0; FY[i] = 0; TQ[i] = 0;
  }
  for (int it = 0; it < w->I; it++) {
    for (int k = 0; k < w->nc; k++) {
      OrCons *c = &w->cons[k];
      if (c->broken) continue;
      if (c->type == OR_ANGLE) solve_angle(w, c, it == 0);
      else solve_pin(w, c, it == 0);
    }
    or_solve_collisions(w);
  }
  for (int i = 0; i < n; i++) {
    VX[i] = (PPX[i] - PX[i]) / h;
    VY[i] = (PPY[i] - PY[i]) / h;
    AV[i] = or_angle_delta(ANGP[i], ANG[i]) / h;
    double sp2 = or_js_hypot(VX[i], VY[i]);
    if (sp2 > w->maxSpeed) { double k = w->maxSpeed / sp2; VX[i] *= k; VY[i] *= k; }
    if (AV[i] > w->maxAngSpeed) AV[i] = w->maxAngSpeed;
    else if (AV[i] < -w->maxAngSpeed) AV[i] = -w->maxAngSpeed;
    PX[i] = PPX[i]; PY[i] = PPY[i];
    ANG[i] = ANGP[i];
    if (ANG[i] > 50 || ANG[i] < -50) { double v = or_wrap_angle(ANG[i]); ANG[i] = v; ANGP[i] = v; }
  }
  measure(w);
}

/* xpbd.js:496-504 */
static int cons_creature(const OrWorld *w, const OrCons *c, int *arm) {
  /* grips are the only constraints whose events are logged */
  for (int ci = 0; ci < w->ncr; ci++) {
    const OrCreature *K = &w->cr[ci];
    if (c->a >= K->first_body && c->a < K->first_body + K->nbodies) { *arm = c->owner; return ci; }
  }
  *arm = c->owner;
  return -1;
}
static void grip_break_f64(OrWorld *w) {
  double inv = 1 / w->h;
  for (int k = 0; k < w->nc; k++) {
    OrCons *c = &w->cons[k];
    if (c->type != OR_GRIP || c->broken) continue;
    double f = or_js_hypot(c->lambdaX, c->lambdaY) * inv;
    c->force = f;
    if (f > c->breakForce) {
      c->broken = 1; w->stats_broken++;
      int arm, ci = cons_creature(w, c, &arm);
      log_event(w, ci, arm, 2);
    }
  }
}

void or_substep(OrWorld *w) {
  if (w->f32) or_f32_substep(w);
  else substep_f64(w);
}
void or_step(OrWorld *w) {
  if (w->f32) {
    or_f32_step(w);
  } else {
    for (int s = 0; s < w->S; s++) substep_f64(w);
    grip_break_f64(w);
  }
  w->step_counter++;
}
/* used by the f32 emulator for event logging */
void or_log_break(OrWorld *w, const OrCons *c) {
  int arm, ci = cons_creature(w, c, &arm);
  log_event(w, ci, arm, 2);
}

void or_get_stats(const OrWorld *w, double *o) {
  o[0] = w->stats_maxSpeed; o[1] = w->stats_maxAngSpeed; o[2] = w->stats_nan;
  o[3] = (double)w->stats_collisions; o[4] = w->stats_broken;
}
int or_events(const OrWorld *w, int *out, int max) {
  int n = w->nev < max ? w->nev : max;
  memcpy(out, w->ev, sizeof(int) * 4 * n);
  return w->nev;
}
void or_clear_events(OrWorld *w) { w->nev = 0; }
void or_set_step_counter(OrWorld *w, int s) { w->step_counter = s; }

/* ── worker-thread throughput baseline ───────────────────────────── */
OrWorld *or_world_clone(const OrWorld *p) {
  OrWorld *w = (OrWorld *)malloc(sizeof(OrWorld));
  *w = *p;
  for (int k = 0; k < OR_NARR; k++) {
    w->arr[k] = (double *)malloc(sizeof(double) * (p->cap > 0 ? p->cap : 1));
    memcpy(w->arr[k], p->arr[k], sizeof(double) * p->n);
  }
  w->cons = (OrCons *)malloc(sizeof(OrCons) * (p->ccap > 0 ? p->ccap : 1));
  memcpy(w->cons, p->cons, sizeof(OrCons) * p->nc);
  w->pos_of_id = (int *)malloc(sizeof(int) * (p->ccap > 0 ? p->ccap : 1));
  memcpy(w->pos_of_id, p->pos_of_id, sizeof(int) * p->nc);
  w->pts = (OrPt *)malloc(sizeof(OrPt) * (p->pcap > 0 ? p->pcap : 1));
  memcpy(w->pts, p->pts, sizeof(OrPt) * p->np);
  w->rects = (double *)malloc(sizeof(double) * 4 * (p->nr > 0 ? p->nr : 1));
  memcpy(w->rects, p->rects, sizeof(double) * 4 * p->nr);
  w->cr = (OrCreature *)malloc(sizeof(OrCreature) * (p->crcap > 0 ? p->crcap : 1));
  memcpy(w->cr, p->cr, sizeof(OrCreature) * p->ncr);
  w->ev = NULL; w->nev = 0; w->evcap = 0;
  w->f32 = NULL;
  return w;
}
typedef struct { OrWorld *w; int steps; uint32_t seed; } BenchArg;
static void *bench_worker(void *p) {
  BenchArg *a = (BenchArg *)p;
  for (int s = 0; s < a->steps; s++) {
    or_apply_law(a->w, a->seed, (uint32_t)s);
    or_step(a->w);
    a->w->nev = 0;
  }
  return NULL;
}
double or_bench_threads(const OrWorld *proto, int threads, int steps, uint32_t seed) {
  if (threads < 1) threads = 1;
  pthread_t *tid = (pthread_t *)malloc(sizeof(pthread_t) * threads);
  BenchArg *args = (BenchArg *)malloc(sizeof(BenchArg) * threads);
  for (int t = 0; t < threads; t++) { args[t].w = or_world_clone(proto); args[t].steps = steps; args[t].seed = seed + t; }
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  for (int t = 0; t < threads; t++) pthread_create(&tid[t], NULL, bench_worker, &args[t]);
  for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  double secs = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
  for (int t = 0; t < threads; t++) or_world_free(args[t].w);
  free(tid); free(args);
  return (double)threads * proto->ncr * steps / secs;
}
