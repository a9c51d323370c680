// Device math shared by the plain and persistent paths.
//
// Numerics, chosen to stay close to the f64 reference (src/physics/xpbd.js)
// in binary32:
//   * angle differences are formed in f32 (a correctly rounded a-b) and only
//     the rare wrap past +-pi is folded in f64 with the exact 2*pi, so wraps
//     carry no f32-TAU error (xpbd.js:128-143);
//   * sin/cos use the accurate library sincosf (angles reach +-50 rad before
//     the finalize wrap, xpbd.js:324);
//   * the persistent path works in creature-local coordinates (integer
//     origin), which keeps position ulps ~8x finer than global f32.
#pragma once
#include "kx_internal.h"

#define KX_SLACK_POS 2e-3f
#define KX_SLACK_POS2 4e-6f
#define KX_SLACK_ANG 2e-4f
#define KX_PI_D 3.141592653589793
#define KX_TAU_D 6.283185307179586
// smallest binary32 above pi: |a-b| rounded below this is certainly < pi
#define KX_PI_BELOW 3.14159250f

__device__ __forceinline__ float kx_fold_d(double x) {
  double r = x;
  if (!(r > -KX_TAU_D && r < KX_TAU_D)) {
    if (!isfinite(r)) return 0.0f;
    r = fmod(r, KX_TAU_D);
  }
  if (r > KX_PI_D) r -= KX_TAU_D;
  else if (r < -KX_PI_D) r += KX_TAU_D;
  return (float)r;
}

// angleDelta(a, b) (xpbd.js:137-143)
__device__ __forceinline__ float kx_angle_delta(float a, float b) {
  float d = a - b;
  if (fabsf(d) <= KX_PI_BELOW) return d;
  return kx_fold_d((double)a - (double)b);
}
// wrapAngle (xpbd.js:128-134), only reached when |a| > 50
__device__ __forceinline__ float kx_wrap_angle(float a) { return kx_fold_d((double)a); }

// ── SDF over the rect list (map.js:91-118), fold order and the best²
//    exterior test preserved. (x, y) are relative to the origin (ox, oy).
struct KxHit { float d, nx, ny; };

__device__ __forceinline__ void kx_fold_rect(float x, float y, float4 r, float ox, float oy,
                                             float &best, float &bnx, float &bny) {
  float rx = r.x - ox, ry = r.y - oy;          // exact for an integer origin
  float rxw = rx + r.z, ryh = ry + r.w;
  float cx = fminf(fmaxf(x, rx), rxw);
  float cy = fminf(fmaxf(y, ry), ryh);
  float dx = x - cx, dy = y - cy;
  float d2 = dx * dx + dy * dy;
  if (d2 > 1e-12f) {
    if (d2 < best * best) {
      float d = sqrtf(d2);
      best = d; bnx = dx / d; bny = dy / d;
    }
  } else {
    float dl = x - rx, dr = rxw - x, dt = y - ry, db = ryh - y;
    float m = dl, nx = -1.0f, ny = 0.0f;
    if (dr < m) { m = dr; nx = 1.0f; ny = 0.0f; }
    if (dt < m) { m = dt; nx = 0.0f; ny = -1.0f; }
    if (db < m) { m = db; nx = 0.0f; ny = 1.0f; }
    if (-m < best) { best = -m; bnx = nx; bny = ny; }
  }
}

// Linear scan (the WGSL shape, solver.wgsl:114-147).
__device__ __forceinline__ bool kx_sdf_linear(const KxDev &d, float x, float y, float ox, float oy, KxHit &h) {
  float best = INFINITY, bnx = 0.0f, bny = 0.0f;
  for (int i = 0; i < d.nr; i++) kx_fold_rect(x, y, d.rects[i], ox, oy, best, bnx, bny);
  if (!isfinite(best)) return false;
  h.d = best; h.nx = bnx; h.ny = bny;
  return true;
}

// Grid-culled fold: only rects registered in the query's cell, ascending.
// Every rect within the cull radius R of any point of a cell is registered in
// it; R >= max(max rect half-thickness, max collision radius) + margin.  Once
// a rect within R is folded |best| <= R and a farther rect can never win
// (d2 < best² or -m < best), and the first near rect overwrites the whole
// fold state, so the result equals the full fold whenever a contact is
// possible; with no rect within R there is no contact either way.
__device__ __forceinline__ bool kx_sdf_grid(const KxDev &d, float x, float y, float ox, float oy, KxHit &h) {
  float gx = (x + ox - d.gx0) * d.inv_cs;
  float gy = (y + oy - d.gy0) * d.inv_cs;
  if (!(gx >= 0.0f && gy >= 0.0f && gx < (float)d.gnx && gy < (float)d.gny)) return false;
  int ix = min((int)gx, d.gnx - 1), iy = min((int)gy, d.gny - 1);
  int cell = iy * d.gnx + ix;
  int b = d.cell_off[cell], e = d.cell_off[cell + 1];
  if (b == e) return false;
  float best = INFINITY, bnx = 0.0f, bny = 0.0f;
  for (int k = b; k < e; k++) kx_fold_rect(x, y, d.rects[d.cell_ids[k]], ox, oy, best, bnx, bny);
  if (!isfinite(best)) return false;
  h.d = best; h.nx = bnx; h.ny = bny;
  return true;
}

// ── constraint projections (xpbd.js:331-439) ─────────────────────────
// Pin / grip: C = (pA + R(thA) la) - (pB + R(thB) lb) or the world anchor,
// exact 2x2 effective mass.  P = float* to global or LDS arrays.
template <class P, class Q>
__device__ __forceinline__ void kx_solve_pin(P ppx, P ppy, P angP, Q im, Q ii, int A, int B,
                                             float lax, float lay, float lbx, float lby,
                                             float alpha, float &lamx, float &lamy) {
  float sa, ca;
  sincosf(angP[A], &sa, &ca);
  float rax = lax * ca - lay * sa, ray = lax * sa + lay * ca;
  float gax = ppx[A] + rax, gay = ppy[A] + ray;
  float gbx = lbx, gby = lby, rbx = 0.0f, rby = 0.0f, wB = 0.0f, wIB = 0.0f;
  if (B >= 0) {
    float sb, cb;
    sincosf(angP[B], &sb, &cb);
    rbx = lbx * cb - lby * sb;
    rby = lbx * sb + lby * cb;
    gbx = ppx[B] + rbx; gby = ppy[B] + rby;
    wB = im[B]; wIB = ii[B];
  }
  float Cx = gax - gbx, Cy = gay - gby;
  if (!isfinite(Cx) || !isfinite(Cy)) return;
  if (Cx * Cx + Cy * Cy < KX_SLACK_POS2) return;
  float pax = -ray, pay = rax, pbx = -rby, pby = rbx;
  float wA = im[A], wIA = ii[A];
  float wSum = wA + wB;
  float k00 = wSum + wIA * pax * pax + wIB * pbx * pbx;
  float k01 = wIA * pax * pay + wIB * pbx * pby;
  float k11 = wSum + wIA * pay * pay + wIB * pby * pby;
  float det = k00 * k11 - k01 * k01;
  if (fabsf(det) < 1e-12f) return;
  float bx = Cx + alpha * lamx;
  float by = Cy + alpha * lamy;
  float inv = 1.0f / det;
  float dlx = -(k11 * inv * bx - k01 * inv * by);
  float dly = -(-k01 * inv * bx + k00 * inv * by);
  if (!isfinite(dlx) || !isfinite(dly)) return;
  lamx += dlx; lamy += dly;
  ppx[A] += wA * dlx; ppy[A] += wA * dly;
  angP[A] += wIA * (pax * dlx + pay * dly);
  if (B >= 0) {
    ppx[B] -= wB * dlx; ppy[B] -= wB * dly;
    angP[B] -= wIB * (pbx * dlx + pby * dly);
  }
}

// Angle motor: scalar XPBD on the deviation from rest, then the one-sided
// joint limit (xpbd.js:392-439).
template <class P, class Q>
__device__ __forceinline__ void kx_solve_angle(P angP, Q ii, int A, int B, float rest, float target,
                                               float mn, float mx, float alpha, float &lam) {
  float wA = ii[A], wB = B >= 0 ? ii[B] : 0.0f;
  float wSum = wA + wB;
  if (wSum <= 0.0f) return;
  float rel = B >= 0 ? kx_angle_delta(angP[B], angP[A]) : angP[A];
  float dev = kx_angle_delta(rel, rest);
  float C = kx_angle_delta(dev, target);
  if (fabsf(C) >= KX_SLACK_ANG) {
    float dl = -(C + alpha * lam) / (wSum + alpha);
    if (!isfinite(dl)) return;
    lam += dl;
    if (B >= 0) { angP[B] += wB * dl; angP[A] -= wA * dl; }
    else angP[A] += wA * dl;
  }
  float rel2 = B >= 0 ? kx_angle_delta(angP[B], angP[A]) : angP[A];
  float dev2 = kx_angle_delta(rel2, rest);
  float over = 0.0f;
  if (dev2 > mx) over = dev2 - mx;
  else if (dev2 < mn) over = dev2 - mn;
  if (fabsf(over) >= KX_SLACK_ANG) {
    float dl2 = -over / wSum;
    if (B >= 0) { angP[B] += wB * dl2; angP[A] -= wA * dl2; }
    else angP[A] += wA * dl2;
  }
}

// Integration of one body (xpbd.js:262-287).
__device__ __forceinline__ void kx_integrate(const KxDev &d, float &vx, float &vy, float &av,
                                             float px, float py, float ang, float im, float ii,
                                             float fx, float fy, float tq,
                                             float &ppx, float &ppy, float &angP) {
  if (im == 0.0f && ii == 0.0f) {
    ppx = px; ppy = py; angP = ang; vx = 0.0f; vy = 0.0f; av = 0.0f;
    return;
  }
  vx = (vx + d.h * im * fx) * d.linDamp;
  vy = (vy + d.h * im * fy) * d.linDamp;
  av = (av + d.h * ii * tq) * d.angDamp;
  float sp = sqrtf(vx * vx + vy * vy);
  if (sp > d.maxSpeed) { float k = d.maxSpeed / sp; vx *= k; vy *= k; }
  if (av > d.maxAngSpeed) av = d.maxAngSpeed;
  else if (av < -d.maxAngSpeed) av = -d.maxAngSpeed;
  if (!isfinite(vx) || !isfinite(vy)) { vx = 0.0f; vy = 0.0f; }
  if (!isfinite(av)) av = 0.0f;
  ppx = px + d.h * vx;
  ppy = py + d.h * vy;
  angP = ang + d.h * av;
}

// Velocity from the position change, re-clamp, commit (xpbd.js:306-325).
__device__ __forceinline__ void kx_finalize(const KxDev &d, float &px, float &py, float &ang,
                                            float &vx, float &vy, float &av,
                                            float ppx, float ppy, float &angP) {
  vx = (ppx - px) * d.invH;
  vy = (ppy - py) * d.invH;
  av = kx_angle_delta(angP, ang) * d.invH;
  float sp = sqrtf(vx * vx + vy * vy);
  if (sp > d.maxSpeed) { float k = d.maxSpeed / sp; vx *= k; vy *= k; }
  if (av > d.maxAngSpeed) av = d.maxAngSpeed;
  else if (av < -d.maxAngSpeed) av = -d.maxAngSpeed;
  px = ppx; py = ppy; ang = angP;
  if (ang > 50.0f || ang < -50.0f) { float w = kx_wrap_angle(ang); ang = w; angP = w; }
}

// ── seeded control law (counter-based mulberry32 chain) ──────────────
__device__ __forceinline__ uint32_t kx_mb32(uint32_t a) {
  a += 0x6D2B79F5u;
  uint32_t t = a;
  t = (t ^ (t >> 15)) * (t | 1u);
  t ^= t + (t ^ (t >> 7)) * (t | 61u);
  return t ^ (t >> 14);
}
__device__ __forceinline__ double kx_ctl_u(uint32_t seed, uint32_t c, uint32_t step, uint32_t slot) {
  uint32_t h = kx_mb32(seed);
  h = kx_mb32(h ^ c);
  h = kx_mb32(h ^ step);
  h = kx_mb32(h ^ slot);
  return (double)h * (1.0 / 4294967296.0);
}
__device__ __forceinline__ double kx_jsmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double kx_jsmax(double a, double b) { return a > b ? a : b; }

// target of a motor under the law: (2u-1)*limit, clamped (core.js:195-198)
__device__ __forceinline__ float kx_law_target(const KxLaw &L, uint32_t cid, uint32_t step, int slot) {
  double u = kx_ctl_u(L.seed, cid, step, (uint32_t)slot);
  double lim = (slot & (KX_SLOTS_PER_ARM - 1)) == 0 ? L.shoulder_limit : L.joint_limit;
  double t = __dmul_rn(__dsub_rn(__dmul_rn(2.0, u), 1.0), lim);
  return (float)kx_jsmax(-lim, kx_jsmin(lim, t));
}

// Adhesion reflex of one arm (core.js:199-218). Reads the tip body's f32
// velocity and pose (tip position in global coordinates as f64), updates
// the grip fields and the f64 cooldown. Returns event bits (attach 1,
// release 2) for this arm.
__device__ __forceinline__ uint32_t kx_reflex(const KxLaw &L, double hS, double adh_u, double &cd,
                                              int &broken, float &bf, float &lbx_g, float &lby_g,
                                              float &lamx, float &lamy, float tvx, float tvy,
                                              double tpx, double tpy, float tang, float glax, float glay) {
  double adh = kx_jsmax(0.0, kx_jsmin(1.0, adh_u));
  if (cd > 0.0) cd = __dsub_rn(cd, hS);
  double bfd = __dadd_rn(L.grip_min, __dmul_rn(__dsub_rn(L.grip_max, L.grip_min), adh));
  if (broken) {
    if (adh > L.attach_min && cd <= 0.0) {
      double vx = tvx, vy = tvy;
      double slip = sqrt(__dadd_rn(__dmul_rn(vx, vx), __dmul_rn(vy, vy)));
      if (slip < L.slip_max) {
        double s, c;
        sincos((double)tang, &s, &c);
        double ax = __dadd_rn(tpx, __dsub_rn(__dmul_rn((double)glax, c), __dmul_rn((double)glay, s)));
        double ay = __dadd_rn(tpy, __dadd_rn(__dmul_rn((double)glax, s), __dmul_rn((double)glay, c)));
        lbx_g = (float)ax; lby_g = (float)ay;
        broken = 0; bf = (float)bfd; lamx = 0.0f; lamy = 0.0f;
        return 1u;
      }
      cd = L.cooldown;
    }
    return 0u;
  }
  bf = (float)bfd;
  if (adh < L.release_below) { broken = 1; return 2u; }
  return 0u;
}
