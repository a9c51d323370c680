// Device math shared by the plain and persistent paths.
//
// Numerics, chosen to stay close to the f64 reference (src/physics/xpbd.js)
// in binary32:
//   * angle differences are folded with 2*pi split into two binary32 parts,
//     so a wrap carries no f32-TAU error (xpbd.js:128-143);
//   * sin/cos: exact two-part reduction by 2*pi, then the hardware sine and
//     cosine (about 4e-7 absolute) over the range angles reach (+-50 rad
//     before the finalize wrap, xpbd.js:324);
//   * both paths work in creature-local coordinates (integer origin), which
//     keeps position ulps several times finer than global f32.
#pragma once
#include "kx_internal.h"

#define KX_SLACK_POS 2e-3f
#define KX_SLACK_POS2 4e-6f
#define KX_SLACK_ANG 2e-4f
// smallest binary32 above pi: |a-b| rounded below this is certainly < pi
#define KX_PI_BELOW 3.14159250f

// Angle folding to [-pi, pi]: round-to-nearest multiple of 2*pi split in two
// binary32 parts (hi exact in 24 bits, lo the remainder), so a fold adds no
// more than an ulp of the result; branch-free for the common |d| <= pi.
#define KX_TAU_HI 6.28318548202514648f
#define KX_TAU_LO -1.74845553e-7f
#define KX_INV_TAU 0.159154943091895336f

__device__ __forceinline__ float kx_fold(float d) {
  // for |d| <= pi, d/(2 pi) rounds to 0 and r == d exactly, so no separate
  // in-range test is needed
  float q = rintf(d * KX_INV_TAU);
  float r = fmaf(-q, KX_TAU_HI, d);
  r = fmaf(-q, KX_TAU_LO, r);
  return isfinite(d) ? r : 0.0f;
}

// kx_fold of a difference known to be finite (a fold result less a finite
// constant): the same operations without the finiteness select
__device__ __forceinline__ float kx_fold_finite(float d) {
  float q = rintf(d * KX_INV_TAU);
  float r = fmaf(-q, KX_TAU_HI, d);
  return fmaf(-q, KX_TAU_LO, r);
}

// angleDelta(a, b) (xpbd.js:137-143): the difference folded into (-pi, pi]
__device__ __forceinline__ float kx_angle_delta(float a, float b) { return kx_fold(a - b); }
// wrapAngle (xpbd.js:128-134), only reached when |a| > 50
__device__ __forceinline__ float kx_wrap_angle(float a) { return kx_fold(a); }

// sin and cos of one argument: Cody-Waite reduction to [-pi, pi] by 2*pi
// split in two binary32 parts with fused multiply-adds, then the
// hardware sine/cosine of the reduced angle in revolutions (v_sin_f32 /
// v_cos_f32). Absolute error about 4e-7 over the angles bodies reach (the
// finalize wrap keeps them within +-50 rad); a 20 px anchor arm is placed to
// within 1e-5 px.
__device__ __forceinline__ void kx_sincos(float x, float *s, float *c) {
  float k = rintf(x * KX_INV_TAU);
  float r = fmaf(-k, KX_TAU_HI, x);
  r = fmaf(-k, KX_TAU_LO, r);
  const float u = r * KX_INV_TAU;
  *s = __builtin_amdgcn_sinf(u);
  *c = __builtin_amdgcn_cosf(u);
}

// 1/x: hardware reciprocal refined by one Newton step
__device__ __forceinline__ float kx_rcp(float x) {
  float r = __builtin_amdgcn_rcpf(x);
  return r * fmaf(-x, r, 2.0f);
}

// ── SDF over the rect list (map.js:91-118), fold order and the best²
//    exterior test preserved. (x, y) are relative to the origin (ox, oy).
//    The fold keeps the winning offset (ux, uy) and whether it is exterior;
//    the unit normal (offset / distance for an exterior hit, the face normal
//    inside) is formed only where a correction uses it.
struct KxHit { float d, ux, uy; bool ext; };

__device__ __forceinline__ void kx_hit_normal(const KxHit &h, float &nx, float &ny) {
  if (h.ext) { nx = h.ux / h.d; ny = h.uy / h.d; }
  else { nx = h.ux; ny = h.uy; }
}

__device__ __forceinline__ void kx_fold_rect(float x, float y, float4 r, float ox, float oy, KxHit &b) {
  float rx = r.x - ox, ry = r.y - oy;          // exact for an integer origin
  float rxw = rx + r.z, ryh = ry + r.w;
  float cx = fminf(fmaxf(x, rx), rxw);
  float cy = fminf(fmaxf(y, ry), ryh);
  float dx = x - cx, dy = y - cy;
  float d2 = dx * dx + dy * dy;
  if (d2 > 1e-12f) {
    if (d2 < b.d * b.d) { b.d = sqrtf(d2); b.ux = dx; b.uy = dy; b.ext = true; }
  } else {
    float dl = x - rx, dr = rxw - x, dt = y - ry, db = ryh - y;
    float m = dl, nx = -1.0f, ny = 0.0f;
    if (dr < m) { m = dr; nx = 1.0f; ny = 0.0f; }
    if (dt < m) { m = dt; nx = 0.0f; ny = -1.0f; }
    if (db < m) { m = db; nx = 0.0f; ny = 1.0f; }
    if (-m < b.d) { b.d = -m; b.ux = nx; b.uy = ny; b.ext = false; }
  }
}

__device__ __forceinline__ void kx_hit_init(KxHit &h) { h.d = INFINITY; h.ux = 0.0f; h.uy = 0.0f; h.ext = false; }

// A creature's integer origin for a step: its first body's coordinate
// rounded; 0 when that coordinate is not finite, so a non-finite body stays
// alone (the reference's solves skip it) instead of shifting its whole
// creature by NaN.
__device__ __forceinline__ float kx_origin(float v) { return isfinite(v) ? rintf(v) : 0.0f; }

// World-frame offset of a body-local point: (x, y) + R(angle) (lx, ly).
__device__ __forceinline__ void kx_world_point(float x, float y, float ca, float sa, float lx, float ly, float &wx,
                                               float &wy) {
  wx = x + (lx * ca - ly * sa);
  wy = y + (lx * sa + ly * ca);
}

// Linear scan (the WGSL shape, solver.wgsl:114-147).
__device__ __forceinline__ bool kx_sdf_linear(const KxDev &d, float x, float y, float ox, float oy, KxHit &h) {
  kx_hit_init(h);
  for (int i = 0; i < d.nr; i++) kx_fold_rect(x, y, d.rects[i], ox, oy, h);
  return isfinite(h.d);
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
  kx_hit_init(h);
  for (int k = b; k < e; k++) kx_fold_rect(x, y, d.cell_rects[k], ox, oy, h);
  return isfinite(h.d);
}

// Same fold, after a test of the cell's bit in an occupancy bitmap held in
// LDS: queries in empty cells (no rect within the cull radius) return at once
// without touching global memory.
__device__ __forceinline__ bool kx_sdf_grid_bm(const KxDev &d, const uint32_t *bits, float x, float y, float ox,
                                               float oy, KxHit &h) {
  float gx = (x + ox - d.gx0) * d.inv_cs;
  float gy = (y + oy - d.gy0) * d.inv_cs;
  if (!(gx >= 0.0f && gy >= 0.0f && gx < (float)d.gnx && gy < (float)d.gny)) return false;
  int ix = min((int)gx, d.gnx - 1), iy = min((int)gy, d.gny - 1);
  int cell = iy * d.gnx + ix;
  if (!((bits[cell >> 5] >> (cell & 31)) & 1u)) return false;
  int b = d.cell_off[cell], e = d.cell_off[cell + 1];
  kx_hit_init(h);
  for (int k = b; k < e; k++) kx_fold_rect(x, y, d.cell_rects[k], ox, oy, h);
  return isfinite(h.d);
}

// ── constraint projections (xpbd.js:331-439) ─────────────────────────
// Pin / grip on values (xpbd.js:331-390): 2x2 XPBD solve on the world
// anchor gap with generalized inverse masses, lambda accumulated per axis.
// (sa, ca) and (sb, cb) are sin/cos of tA and tB. Branch-free: every lane
// evaluates the whole solve and writes the projected states; the return value
// says whether they apply (gap beyond the dead zone, regular and finite
// solve) and callers store them only then.
__device__ __forceinline__ bool kx_pin_core_sc(float &xA, float &yA, float &tA, float wA, float wIA, float &xB,
                                               float &yB, float &tB, float wB, float wIB, bool hasB, float lax,
                                               float lay, float lbx, float lby, float alpha, float &lamx,
                                               float &lamy, float sa, float ca, float sb, float cb) {
  float rax = lax * ca - lay * sa, ray = lax * sa + lay * ca;
  float gax = xA + rax, gay = yA + ray;
  float tbx = lbx * cb - lby * sb, tby = lbx * sb + lby * cb;
  float rbx = hasB ? tbx : 0.0f, rby = hasB ? tby : 0.0f;
  float gbx = hasB ? xB + tbx : lbx, gby = hasB ? yB + tby : lby;
  if (!hasB) { wB = 0.0f; wIB = 0.0f; }
  float Cx = gax - gbx, Cy = gay - gby;
  // a non-finite C makes dlx/dly non-finite and is rejected below
  bool ok = !(Cx * Cx + Cy * Cy < KX_SLACK_POS2);
  float pax = -ray, pay = rax, pbx = -rby, pby = rbx;
  float wSum = wA + wB;
  float k00 = wSum + wIA * pax * pax + wIB * pbx * pbx;
  float k01 = wIA * pax * pay + wIB * pbx * pby;
  float k11 = wSum + wIA * pay * pay + wIB * pby * pby;
  float det = k00 * k11 - k01 * k01;
  ok = ok && !(fabsf(det) < 1e-12f);
  float bx = Cx + alpha * lamx;
  float by = Cy + alpha * lamy;
  float inv = kx_rcp(det);
  float dlx = -(k11 * inv * bx - k01 * inv * by);
  float dly = -(-k01 * inv * bx + k00 * inv * by);
  ok = ok && isfinite(dlx) && isfinite(dly);
  lamx += dlx; lamy += dly;
  xA += wA * dlx; yA += wA * dly;
  tA += wIA * (pax * dlx + pay * dly);
  xB -= wB * dlx; yB -= wB * dly;
  tB -= wIB * (pbx * dlx + pby * dly);
  return ok;
}

__device__ __forceinline__ bool kx_pin_core(float &xA, float &yA, float &tA, float wA, float wIA, float &xB,
                                            float &yB, float &tB, float wB, float wIB, bool hasB, float lax,
                                            float lay, float lbx, float lby, float alpha, float &lamx,
                                            float &lamy) {
  float sa, ca, sb, cb;
  kx_sincos(tA, &sa, &ca);
  kx_sincos(tB, &sb, &cb);
  return kx_pin_core_sc(xA, yA, tA, wA, wIA, xB, yB, tB, wB, wIB, hasB, lax, lay, lbx, lby, alpha, lamx, lamy, sa,
                        ca, sb, cb);
}

// Angle motor on values: scalar XPBD on the deviation from rest, then the
// one-sided joint limit (xpbd.js:392-439). rS is kx_rcp(wA + wB) (wB = 0
// without a second body), a constant of the constraint that callers may
// compute once. Returns true if an angle moved.
__device__ __forceinline__ bool kx_angle_core(float &tA, float &tB, float wA, float wB, bool hasB, float rest,
                                              float target, float mn, float mx, float alpha, float rS, float &lam) {
  if (!hasB) wB = 0.0f;
  float wSum = wA + wB;
  // both sides evaluated, then selected: no branch on hasB
  const float dBA = kx_angle_delta(tB, tA);
  float rel = hasB ? dBA : tA;
  float dev = kx_angle_delta(rel, rest);
  // angleDelta(dev, target) without its finiteness select: dev is finite,
  // so C is non-finite only for a non-finite target, where the fold gives
  // NaN instead of 0; both fail the slack test below, and C is used only
  // when that test passes
  float C = kx_fold_finite(dev - target);
  const bool live = wSum > 0.0f;
  const bool corr = fabsf(C) >= KX_SLACK_ANG;
  float dl = -(C + alpha * lam) * kx_rcp(wSum + alpha);
  // a non-finite correction rejects the whole constraint, limit included
  const bool bad = corr && !isfinite(dl);
  const bool ok1 = live && corr && !bad;
  float lam1 = lam + dl;
  float tA1 = hasB ? tA - wA * dl : tA + wA * dl;
  float tB1 = tB + wB * dl;
  lam = ok1 ? lam1 : lam;
  tA = ok1 ? tA1 : tA;
  tB = (ok1 && hasB) ? tB1 : tB;
  const float dBA2 = kx_angle_delta(tB, tA);
  float rel2 = hasB ? dBA2 : tA;
  float dev2 = kx_angle_delta(rel2, rest);
  float over = dev2 > mx ? dev2 - mx : (dev2 < mn ? dev2 - mn : 0.0f);
  const bool ok2 = live && !bad && fabsf(over) >= KX_SLACK_ANG;
  float dl2 = -over * rS;
  float tA2 = hasB ? tA - wA * dl2 : tA + wA * dl2;
  float tB2 = tB + wB * dl2;
  tA = ok2 ? tA2 : tA;
  tB = (ok2 && hasB) ? tB2 : tB;
  return ok1 || ok2;
}

// kx_angle_core for a motor between two bodies with a finite rest angle and
// limits mn <= mx (or NaN): the operations of its hasB path in the same
// order, without the selects of a world-anchored motor, and without the
// finiteness select of dev and dev2, whose arguments (a fold result less the
// finite rest) are always finite. The limit excess is max(dev2 - mx, 0) +
// min(dev2 - mn, 0): with mn <= mx at most one term is non-zero and the other
// adds an exact zero, so it equals the nested select (a NaN limit drops its
// term in both forms). rM is kx_rcp((wA + wB) + alpha), which a caller holding
// alpha fixed over many solves computes once. LIVE: the caller has already
// required wA + wB > 0 of every result it commits.
template <bool LIVE = false>
__device__ __forceinline__ bool kx_angle_core_ab(float &tA, float &tB, float wA, float wB, float rest, float target,
                                                 float mn, float mx, float alpha, float rS, float rM, float &lam) {
  const float wSum = wA + wB;
  const float rel = kx_angle_delta(tB, tA);
  const float dev = kx_fold_finite(rel - rest);
  const float C = kx_fold_finite(dev - target);   // as in kx_angle_core
  const bool live = LIVE || wSum > 0.0f;
  const bool corr = fabsf(C) >= KX_SLACK_ANG;
  float dl = -(C + alpha * lam) * rM;
  const bool bad = corr && !isfinite(dl);
  const bool ok1 = live && corr && !bad;
  float lam1 = lam + dl;
  float tA1 = tA - wA * dl;
  float tB1 = tB + wB * dl;
  lam = ok1 ? lam1 : lam;
  tA = ok1 ? tA1 : tA;
  tB = ok1 ? tB1 : tB;
  const float rel2 = kx_angle_delta(tB, tA);
  const float dev2 = kx_fold_finite(rel2 - rest);
  float over = fmaxf(dev2 - mx, 0.0f) + fminf(dev2 - mn, 0.0f);
  const bool ok2 = live && !bad && fabsf(over) >= KX_SLACK_ANG;
  float dl2 = -over * rS;
  float tA2 = tA - wA * dl2;
  float tB2 = tB + wB * dl2;
  tA = ok2 ? tA2 : tA;
  tB = ok2 ? tB2 : tB;
  return ok1 || ok2;
}

// SoA forms used by the plain path (P, Q = float* to global arrays).
template <class P, class Q>
__device__ __forceinline__ void kx_solve_pin(P ppx, P ppy, P angP, Q im, Q ii, int A, int B, float lax, float lay,
                                             float lbx, float lby, float alpha, float &lamx, float &lamy) {
  float xA = ppx[A], yA = ppy[A], tA = angP[A];
  float xB = 0.0f, yB = 0.0f, tB = 0.0f, wB = 0.0f, wIB = 0.0f;
  if (B >= 0) { xB = ppx[B]; yB = ppy[B]; tB = angP[B]; wB = im[B]; wIB = ii[B]; }
  float nlx = lamx, nly = lamy;
  if (!kx_pin_core(xA, yA, tA, im[A], ii[A], xB, yB, tB, wB, wIB, B >= 0, lax, lay, lbx, lby, alpha, nlx, nly))
    return;
  lamx = nlx; lamy = nly;
  ppx[A] = xA; ppy[A] = yA; angP[A] = tA;
  if (B >= 0) { ppx[B] = xB; ppy[B] = yB; angP[B] = tB; }
}

template <class P, class Q>
__device__ __forceinline__ void kx_solve_angle(P angP, Q ii, int A, int B, float rest, float target, float mn,
                                               float mx, float alpha, float &lam) {
  float tA = angP[A], tB = B >= 0 ? angP[B] : 0.0f;
  const float wA = ii[A], wB = B >= 0 ? ii[B] : 0.0f;
  if (!kx_angle_core(tA, tB, wA, wB, B >= 0, rest, target, mn, mx, alpha, kx_rcp(wA + wB), lam)) return;
  angP[A] = tA;
  if (B >= 0) angP[B] = tB;
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
  // sqrtf(q) > maxSpeed as q >= maxSpeed2 (the square root only to clamp)
  const float q = vx * vx + vy * vy;
  if (q >= d.maxSpeed2) { float k = d.maxSpeed / sqrtf(q); vx *= k; vy *= k; }
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
  const float q = vx * vx + vy * vy;   // as in kx_integrate
  if (q >= d.maxSpeed2) { float k = d.maxSpeed / sqrtf(q); vx *= k; vy *= k; }
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
// the chain's state after seed, creature and step (shared by every slot of
// a creature's step), and the draw of one slot from it
__device__ __forceinline__ uint32_t kx_ctl_prefix(uint32_t seed, uint32_t c, uint32_t step) {
  return kx_mb32(kx_mb32(kx_mb32(seed) ^ c) ^ step);
}
__device__ __forceinline__ double kx_ctl_u_at(uint32_t prefix, uint32_t slot) {
  return (double)kx_mb32(prefix ^ slot) * (1.0 / 4294967296.0);
}
__device__ __forceinline__ double kx_ctl_u(uint32_t seed, uint32_t c, uint32_t step, uint32_t slot) {
  return kx_ctl_u_at(kx_ctl_prefix(seed, c, step), slot);
}
__device__ __forceinline__ double kx_jsmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ double kx_jsmax(double a, double b) { return a > b ? a : b; }

// target of a motor under the law: (2u-1)*limit, clamped (core.js:195-198)
__device__ __forceinline__ float kx_law_target_at(const KxLaw &L, uint32_t prefix, int slot) {
  double u = kx_ctl_u_at(prefix, (uint32_t)slot);
  double lim = (slot & (KX_SLOTS_PER_ARM - 1)) == 0 ? L.shoulder_limit : L.joint_limit;
  double t = __dmul_rn(__dsub_rn(__dmul_rn(2.0, u), 1.0), lim);
  return (float)kx_jsmax(-lim, kx_jsmin(lim, t));
}
__device__ __forceinline__ float kx_law_target(const KxLaw &L, uint32_t cid, uint32_t step, int slot) {
  return kx_law_target_at(L, kx_ctl_prefix(L.seed, cid, step), slot);
}

// Control tape: per step and creature the inputs of applyControl
// (core.js:183-198) as decode leaves them (src/brain/net.js:114-124):
// rec[0] tone, rec[1+3a] shoulder, rec[2+3a] curl, rec[3+3a] adhesion of arm a.
__device__ __forceinline__ const float *kx_tape_rec(const KxDev &d, int row, int c) {
  return d.tape + ((size_t)row * d.nCr + c) * KX_TAPE_WORDS;
}
// toneToCompliance of the clamped tone (core.js:185-187)
__device__ __forceinline__ double kx_tape_comp(const KxLaw &L, float tone) {
  double t = kx_jsmax(0.0, kx_jsmin(1.0, (double)tone));
  return __ddiv_rn(1.0, __dadd_rn(L.stiff_min, __dmul_rn(t, __dsub_rn(L.stiff_max, L.stiff_min))));
}
// shoulder target clamped to the shoulder limit, every joint of the arm the
// curl clamped to the joint limit (core.js:193-197)
__device__ __forceinline__ float kx_tape_target(const KxLaw &L, const float *rec, int slot) {
  const int arm = slot / KX_SLOTS_PER_ARM, m = slot % KX_SLOTS_PER_ARM;
  const double lim = m == 0 ? L.shoulder_limit : L.joint_limit;
  const double u = (double)rec[(m == 0 ? 1 : 2) + 3 * arm];
  return (float)kx_jsmax(-lim, kx_jsmin(lim, u));
}

// Adhesion reflex of one arm (core.js:199-218). Reads the tip body's f32
// velocity and pose (tip position in global coordinates as f64), updates
// the grip fields and the f64 cooldown. Returns event bits (attach 1,
// release 2) for this arm.
__device__ __forceinline__ uint32_t kx_reflex(const KxLaw &L, double hS, double adh_u, double grip_u, double &cd,
                                              int &broken, float &bf, float &lbx_g, float &lby_g,
                                              float &lamx, float &lamy, float tvx, float tvy,
                                              double tpx, double tpy, float tang, float glax, float glay) {
  double adh = kx_jsmax(0.0, kx_jsmin(1.0, adh_u));
  if (cd > 0.0) cd = __dsub_rn(cd, hS);
  const double ga = (L.flags & KX_LAW_FIXED_GRIP) ? kx_jsmax(0.0, kx_jsmin(1.0, grip_u)) : adh;
  double bfd = __dadd_rn(L.grip_min, __dmul_rn(__dsub_rn(L.grip_max, L.grip_min), ga));
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
