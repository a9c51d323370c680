/*
 * f32 emulation of the WGSL solver — TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Restates tools/f32solver.mjs (f32ify, collMode 'seq'): every arithmetic
 * result rounded to binary32, cos/sin/exp evaluated in binary64 and rounded,
 * statics snapshot once, controls snapshot at the start of every step.
 * Compiled with -ffp-contract=off so no operation is fused.
 */
#include "oracle.h"
#include "oracle_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct OrF32 {
  float h, invH, invH2, linDamp, angDamp, maxSpeed, maxAngSpeed;
  int n, nc, np, nr;
  float *imF, *iiF;
  float *laxF, *layF, *restF, *minF, *maxF;
  int *type, *a, *b;
  float *compF, *tgtF, *lbxF, *lbyF, *brkF;
  float *rx, *ry, *rxw, *ryh;
  int *bodyStart, *bodyCnt, *order;   /* points grouped by body, stable */
};

/* f32solver.mjs:26-33 */
#define SLACK_POS_F ((float)0.002)
#define SLACK_ANG_F ((float)0.0002)
#define TAU_F ((float)6.28318530718)
#define PI_F ((float)3.14159265359)
#define EPS_DET_F ((float)1e-12)
#define EPS_SDF_F ((float)1e-12)
#define SDF_INF_F ((float)1e30)
#define SDF_VALID_F ((float)1e29)

/* f32solver.mjs:36-43: FRem semantics x - TAU*trunc(x/TAU), quotient in f64 */
static float angle_delta_f(float a, float b) {
  float x = a - b;
  double q = trunc((double)x / (double)TAU_F);
  float d = x - (float)((double)TAU_F * q);
  if (d > PI_F) return d - TAU_F;
  if (d < -PI_F) return d + TAU_F;
  return d;
}
static float wrap_angle_f(float a) {
  double q = trunc((double)a / (double)TAU_F);
  float d = a - (float)((double)TAU_F * q);
  if (d > PI_F) return d - TAU_F;
  if (d < -PI_F) return d + TAU_F;
  return d;
}
static float length_f(float x, float y) { return (float)sqrt((double)(x * x + y * y)); }
static float cosf64(float x) { return (float)cos((double)x); }
static float sinf64(float x) { return (float)sin((double)x); }

static float *fa(int n) { return (float *)calloc(n > 0 ? n : 1, sizeof(float)); }
static int *ia(int n) { return (int *)calloc(n > 0 ? n : 1, sizeof(int)); }

void or_f32ify(OrWorld *w) {
  OrF32 *F = (OrF32 *)calloc(1, sizeof(OrF32));
  F->n = w->n; F->nc = w->nc; F->np = w->np; F->nr = w->nr;
  F->h = (float)w->h;
  F->invH = (float)(1 / w->h);
  F->invH2 = (float)(1 / (w->h * w->h));
  F->linDamp = (float)exp((double)(-(float)w->linDrag * F->h));
  F->angDamp = (float)exp((double)(-(float)w->angDrag * F->h));
  F->maxSpeed = (float)w->maxSpeed;
  F->maxAngSpeed = (float)w->maxAngSpeed;
  int n = w->n, nc = w->nc, np = w->np;
  F->imF = fa(n); F->iiF = fa(n);
  for (int i = 0; i < n; i++) { F->imF[i] = (float)IM[i]; F->iiF[i] = (float)II[i]; }
  F->laxF = fa(nc); F->layF = fa(nc); F->restF = fa(nc); F->minF = fa(nc); F->maxF = fa(nc);
  F->type = ia(nc); F->a = ia(nc); F->b = ia(nc);
  F->compF = fa(nc); F->tgtF = fa(nc); F->lbxF = fa(nc); F->lbyF = fa(nc); F->brkF = fa(nc);
  for (int k = 0; k < nc; k++) {
    const OrCons *c = &w->cons[k];
    F->laxF[k] = (float)c->lax; F->layF[k] = (float)c->lay; F->restF[k] = (float)c->rest;
    F->minF[k] = (float)c->minA; F->maxF[k] = (float)c->maxA;
    F->type[k] = c->type; F->a[k] = c->a; F->b[k] = c->b;
  }
  F->rx = fa(w->nr); F->ry = fa(w->nr); F->rxw = fa(w->nr); F->ryh = fa(w->nr);
  for (int i = 0; i < w->nr; i++) {
    float x = (float)w->rects[4 * i], y = (float)w->rects[4 * i + 1];
    float ww = (float)w->rects[4 * i + 2], hh = (float)w->rects[4 * i + 3];
    F->rx[i] = x; F->ry[i] = y; F->rxw[i] = x + ww; F->ryh[i] = y + hh;
  }
  F->bodyStart = ia(n); F->bodyCnt = ia(n); F->order = ia(np);
  for (int i = 0; i < np; i++) F->bodyCnt[w->pts[i].body]++;
  int acc = 0;
  for (int b = 0; b < n; b++) { F->bodyStart[b] = acc; acc += F->bodyCnt[b]; }
  int *cur = ia(n);
  memcpy(cur, F->bodyStart, sizeof(int) * n);
  for (int i = 0; i < np; i++) F->order[cur[w->pts[i].body]++] = i;
  free(cur);
  /* roundState (f32solver.mjs:128-135) */
  for (int i = 0; i < n; i++) {
    PX[i] = (float)PX[i]; PY[i] = (float)PY[i]; PPX[i] = (float)PPX[i]; PPY[i] = (float)PPY[i];
    VX[i] = (float)VX[i]; VY[i] = (float)VY[i]; ANG[i] = (float)ANG[i]; ANGP[i] = (float)ANGP[i];
    AV[i] = (float)AV[i];
  }
  free(w->f32);
  w->f32 = F;
}

static void snapshot_controls(OrWorld *w) {
  OrF32 *F = w->f32;
  for (int k = 0; k < w->nc; k++) {
    const OrCons *c = &w->cons[k];
    F->compF[k] = (float)c->compliance; F->tgtF[k] = (float)c->target;
    F->lbxF[k] = (float)c->lbx; F->lbyF[k] = (float)c->lby;
    F->brkF[k] = isfinite(c->breakForce) ? (float)or_js_min(c->breakForce, 1e30) : (float)1e30;
  }
}

/* f32solver.mjs:139-166 */
static int sdf_f(const OrF32 *F, float x, float y, float *out) {
  float best = SDF_INF_F, bnx = 0, bny = 0;
  for (int i = 0; i < F->nr; i++) {
    float cx = x < F->rx[i] ? F->rx[i] : (x > F->rxw[i] ? F->rxw[i] : x);
    float cy = y < F->ry[i] ? F->ry[i] : (y > F->ryh[i] ? F->ryh[i] : y);
    float dx = x - cx, dy = y - cy;
    float d2 = dx * dx + dy * dy;
    if (d2 > EPS_SDF_F) {
      if (d2 < best * best) {
        best = (float)sqrt((double)d2);
        bnx = dx / best; bny = dy / best;
      }
    } else {
      float dl = x - F->rx[i], dr = F->rxw[i] - x, dt = y - F->ry[i], db = F->ryh[i] - y;
      float m = dl, nx = -1, ny = 0;
      if (dr < m) { m = dr; nx = 1; ny = 0; }
      if (dt < m) { m = dt; nx = 0; ny = -1; }
      if (db < m) { m = db; nx = 0; ny = 1; }
      float d = -m;
      if (d < best) { best = d; bnx = nx; bny = ny; }
    }
  }
  if (best > SDF_VALID_F) return 0;
  out[0] = best; out[1] = bnx; out[2] = bny;
  return 1;
}

static void reset_lambda_f(OrWorld *w) {
  for (int k = 0; k < w->nc; k++) { w->cons[k].lambdaX = 0; w->cons[k].lambdaY = 0; w->cons[k].lambda = 0; }
}

/* f32solver.mjs:177-200 */
static void integrate_f(OrWorld *w) {
  const OrF32 *F = w->f32;
  for (int i = 0; i < w->n; i++) {
    if (F->imF[i] == 0 && F->iiF[i] == 0) {
      PPX[i] = PX[i]; PPY[i] = PY[i]; ANGP[i] = ANG[i];
      VX[i] = 0; VY[i] = 0; AV[i] = 0;
      continue;
    }
    float vx = ((float)VX[i] + (F->h * F->imF[i]) * (float)FX[i]) * F->linDamp;
    float vy = ((float)VY[i] + (F->h * F->imF[i]) * (float)FY[i]) * F->linDamp;
    float av = ((float)AV[i] + (F->h * F->iiF[i]) * (float)TQ[i]) * F->angDamp;
    float sp = length_f(vx, vy);
    if (sp > F->maxSpeed) { float k = F->maxSpeed / sp; vx = vx * k; vy = vy * k; }
    if (av > F->maxAngSpeed) av = F->maxAngSpeed;
    else if (av < -F->maxAngSpeed) av = -F->maxAngSpeed;
    if (!(vx == vx) || !(vy == vy)) { vx = 0; vy = 0; }
    if (!(av == av)) av = 0;
    VX[i] = vx; VY[i] = vy; AV[i] = av;
    PPX[i] = (float)PX[i] + F->h * vx;
    PPY[i] = (float)PY[i] + F->h * vy;
    ANGP[i] = (float)ANG[i] + F->h * av;
    FX[i] = 0; FY[i] = 0; TQ[i] = 0;
  }
}

/* f32solver.mjs:203-256 */
static void solve_pin_f(OrWorld *w, int k) {
  const OrF32 *F = w->f32;
  OrCons *c = &w->cons[k];
  int A = F->a[k], B = F->b[k];
  float ca = cosf64((float)ANGP[A]), sa = sinf64((float)ANGP[A]);
  float rax = F->laxF[k] * ca - F->layF[k] * sa;
  float ray = F->laxF[k] * sa + F->layF[k] * ca;
  float gax = (float)PPX[A] + rax, gay = (float)PPY[A] + ray;
  float gbx = F->lbxF[k], gby = F->lbyF[k];
  float rbx = 0, rby = 0, wB = 0, wIB = 0;
  if (B >= 0) {
    float cb = cosf64((float)ANGP[B]), sb = sinf64((float)ANGP[B]);
    rbx = F->lbxF[k] * cb - F->lbyF[k] * sb;
    rby = F->lbxF[k] * sb + F->lbyF[k] * cb;
    gbx = (float)PPX[B] + rbx; gby = (float)PPY[B] + rby;
    wB = F->imF[B]; wIB = F->iiF[B];
  }
  float Cx = gax - gbx, Cy = gay - gby;
  if (!(Cx == Cx) || !(Cy == Cy)) return;
  if (Cx * Cx + Cy * Cy < SLACK_POS_F * SLACK_POS_F) return;
  float pax = -ray, pay = rax;
  float pbx = -rby, pby = rbx;
  float wA = F->imF[A], wIA = F->iiF[A];
  float wSum = wA + wB;
  float k00 = (wSum + (wIA * pax) * pax) + (wIB * pbx) * pbx;
  float k01 = (wIA * pax) * pay + (wIB * pbx) * pby;
  float k11 = (wSum + (wIA * pay) * pay) + (wIB * pby) * pby;
  float det = k00 * k11 - k01 * k01;
  if (fabsf(det) < EPS_DET_F) return;
  float alphaTilde = F->compF[k] * F->invH2;
  float lx = (float)c->lambdaX, ly = (float)c->lambdaY;
  float bx = Cx + alphaTilde * lx;
  float by = Cy + alphaTilde * ly;
  float i00 = k11 / det, i11 = k00 / det, i01 = -(k01 / det);
  float dlx = -(i00 * bx + i01 * by);
  float dly = -(i01 * bx + i11 * by);
  if (!(dlx == dlx) || !(dly == dly)) return;
  c->lambdaX = lx + dlx;
  c->lambdaY = ly + dly;
  PPX[A] = (float)PPX[A] + wA * dlx;
  PPY[A] = (float)PPY[A] + wA * dly;
  ANGP[A] = (float)ANGP[A] + wIA * (pax * dlx + pay * dly);
  if (B >= 0) {
    PPX[B] = (float)PPX[B] - wB * dlx;
    PPY[B] = (float)PPY[B] - wB * dly;
    ANGP[B] = (float)ANGP[B] - wIB * (pbx * dlx + pby * dly);
  }
}

/* f32solver.mjs:258-301 */
static void solve_angle_f(OrWorld *w, int k) {
  const OrF32 *F = w->f32;
  OrCons *c = &w->cons[k];
  int A = F->a[k], B = F->b[k];
  int hasB = B >= 0;
  float wA = F->iiF[A];
  float wB = 0;
  float rel = (float)ANGP[A];
  if (hasB) { wB = F->iiF[B]; rel = angle_delta_f((float)ANGP[B], (float)ANGP[A]); }
  float wSum = wA + wB;
  if (wSum <= 0) return;
  float dev = angle_delta_f(rel, F->restF[k]);
  float C = angle_delta_f(dev, F->tgtF[k]);
  if (!(C == C)) return;
  if (fabsf(C) >= SLACK_ANG_F) {
    float alphaTilde = F->compF[k] * F->invH2;
    float lam = (float)c->lambda;
    float dl = -((C + alphaTilde * lam) / (wSum + alphaTilde));
    if (!(dl == dl)) return;
    c->lambda = lam + dl;
    if (hasB) {
      ANGP[B] = (float)ANGP[B] + wB * dl;
      ANGP[A] = (float)ANGP[A] - wA * dl;
    } else {
      ANGP[A] = (float)ANGP[A] + wA * dl;
    }
  }
  float rel2 = (float)ANGP[A];
  if (hasB) rel2 = angle_delta_f((float)ANGP[B], (float)ANGP[A]);
  float dev2 = angle_delta_f(rel2, F->restF[k]);
  float over = 0;
  if (dev2 > F->maxF[k]) over = dev2 - F->maxF[k];
  else if (dev2 < F->minF[k]) over = dev2 - F->minF[k];
  if (fabsf(over) >= SLACK_ANG_F) {
    float dl2 = -(over / wSum);
    if (hasB) {
      ANGP[B] = (float)ANGP[B] + wB * dl2;
      ANGP[A] = (float)ANGP[A] - wA * dl2;
    } else {
      ANGP[A] = (float)ANGP[A] + wA * dl2;
    }
  }
}

/* f32solver.mjs:317-344 (collMode 'seq') */
static void solve_collisions_f(OrWorld *w) {
  const OrF32 *F = w->f32;
  float hit[3];
  for (int b = 0; b < w->n; b++) {
    if (F->imF[b] == 0) continue;
    for (int q = 0; q < F->bodyCnt[b]; q++) {
      const OrPt *cp = &w->pts[F->order[F->bodyStart[b] + q]];
      float ca = cosf64((float)ANGP[b]), sa = sinf64((float)ANGP[b]);
      float lx = (float)cp->lx, ly = (float)cp->ly, r = (float)cp->r;
      float wx = (float)PPX[b] + (lx * ca - ly * sa);
      float wy = (float)PPY[b] + (lx * sa + ly * ca);
      if (!sdf_f(F, wx, wy, hit)) continue;
      float pen = r - hit[0];
      if (pen <= SLACK_POS_F) continue;
      PPX[b] = (float)PPX[b] + hit[1] * pen;
      PPY[b] = (float)PPY[b] + hit[2] * pen;
      w->stats_collisions++;
    }
  }
}

/* f32solver.mjs:376-399 */
static void finalize_f(OrWorld *w) {
  const OrF32 *F = w->f32;
  for (int i = 0; i < w->n; i++) {
    if (F->imF[i] == 0 && F->iiF[i] == 0) { VX[i] = 0; VY[i] = 0; AV[i] = 0; continue; }
    float vx = ((float)PPX[i] - (float)PX[i]) * F->invH;
    float vy = ((float)PPY[i] - (float)PY[i]) * F->invH;
    float av = angle_delta_f((float)ANGP[i], (float)ANG[i]) * F->invH;
    float sp = length_f(vx, vy);
    if (sp > F->maxSpeed) { float k = F->maxSpeed / sp; vx = vx * k; vy = vy * k; }
    if (av > F->maxAngSpeed) av = F->maxAngSpeed;
    else if (av < -F->maxAngSpeed) av = -F->maxAngSpeed;
    VX[i] = vx; VY[i] = vy; AV[i] = av;
    PX[i] = PPX[i]; PY[i] = PPY[i];
    ANG[i] = ANGP[i];
    if (ANG[i] > 50 || ANG[i] < -50) { float r = wrap_angle_f((float)ANG[i]); ANG[i] = r; ANGP[i] = r; }
  }
}

/* f32solver.mjs:402-416 */
static void grip_break_f(OrWorld *w) {
  const OrF32 *F = w->f32;
  double invH64 = 1 / w->h;
  for (int k = 0; k < w->nc; k++) {
    OrCons *c = &w->cons[k];
    if (F->type[k] != OR_GRIP || c->broken) continue;
    float fm = length_f((float)c->lambdaX, (float)c->lambdaY) * F->invH;
    if (fm > F->brkF[k]) {
      c->broken = 1; c->lambdaX = 0; c->lambdaY = 0; c->force = 0;
      w->stats_broken++;
      or_log_break(w, c);
    } else {
      c->force = or_js_hypot(c->lambdaX, c->lambdaY) * invH64;
    }
  }
}

static void measure_f(OrWorld *w) {
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

void or_f32_substep(OrWorld *w) {
  reset_lambda_f(w);
  integrate_f(w);
  for (int it = 0; it < w->I; it++) {
    for (int k = 0; k < w->nc; k++) {
      if (w->cons[k].broken) continue;
      if (w->f32->type[k] == OR_ANGLE) solve_angle_f(w, k);
      else solve_pin_f(w, k);
    }
    solve_collisions_f(w);
  }
  finalize_f(w);
  measure_f(w);
}

void or_f32_step(OrWorld *w) {
  snapshot_controls(w);
  for (int s = 0; s < w->S; s++) or_f32_substep(w);
  grip_break_f(w);
}
