// Synthetic world builder for benchmarks and serving requests.
//
// Restates the reference's world generators in f64, in the same operation
// order: makeMap / mulberry32 / findSpawn (src/world/map.js:11-71, 120-128)
// and spawnKraken (src/body/kraken.js:64-148), the latter generalised to
// `arms` x `segs` bodies with a mantle ring of `ring` collision points.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/kxpbd.h"

namespace {
// BODY (kraken.js:22-47)
constexpr double kSegLen = 13.5, kShoulderR = 24, kMantleR = 18, kMantleMass = 20, kMantleIF = 2.5;
constexpr double kMassProx = 3.2, kMassDist = 1.8, kThickProx = 11.9, kThickDist = 5.9;
constexpr double kMuscleMin = 2000;
const double kJointLimit = M_PI * 0.42, kShoulderLimit = M_PI * 0.62;

double rng_next(uint32_t &s) {
  s += 0x6D2B79F5u;
  uint32_t t = s;
  t = (t ^ (t >> 15)) * (t | 1u);
  t ^= t + (t ^ (t >> 7)) * (t | 61u);
  return (double)(t ^ (t >> 14)) / 4294967296.0;
}
double jmin(double a, double b) { return a < b ? a : b; }
double jmax(double a, double b) { return a > b ? a : b; }
}  // namespace

struct kx_builder {
  double map_w = 0, map_h = 0;
  std::vector<double> rects;
  std::vector<double> px, py, ang, im, ii;
  std::vector<int32_t> ctype, ca, cb, cown;
  std::vector<double> lax, lay, lbx, lby, comp, rest, cmin, cmax, target, bf;
  std::vector<uint8_t> broken;
  std::vector<int32_t> pbody;
  std::vector<double> plx, ply, pr;
  std::vector<int32_t> arm_off{0}, arm_tip, arm_grip, motor_off{0}, motors;
  // f32 views handed out by kx_builder_desc
  std::vector<float> f_im, f_ii, f_lax, f_lay, f_lbx, f_lby, f_rest, f_min, f_max, f_plx, f_ply, f_pr, f_rects;

  int add_body(double x, double y, double a, double mass, double inertia) {
    px.push_back(x); py.push_back(y); ang.push_back(a);
    im.push_back(mass > 0 ? 1 / mass : 0);
    ii.push_back(inertia > 0 ? 1 / inertia : 0);
    return (int)px.size() - 1;
  }
  int add_cons(int t, int a, int b, double la_x, double la_y, double lb_x, double lb_y, double compliance) {
    ctype.push_back(t); ca.push_back(a); cb.push_back(b); cown.push_back(-1);
    lax.push_back(la_x); lay.push_back(la_y); lbx.push_back(lb_x); lby.push_back(lb_y);
    comp.push_back(compliance); rest.push_back(0); cmin.push_back(-INFINITY); cmax.push_back(INFINITY);
    target.push_back(0); bf.push_back(INFINITY); broken.push_back(0);
    return (int)ctype.size() - 1;
  }
  void add_point(int b, double x, double y, double r) {
    pbody.push_back(b); plx.push_back(x); ply.push_back(y); pr.push_back(r);
  }
  bool sdf(double x, double y, double &d) const {
    double best = INFINITY;
    size_t n = rects.size() / 4;
    for (size_t i = 0; i < n; i++) {
      double rx = rects[4 * i], ry = rects[4 * i + 1], rw = rects[4 * i + 2], rh = rects[4 * i + 3];
      double cx = jmax(rx, jmin(x, rx + rw)), cy = jmax(ry, jmin(y, ry + rh));
      double dx = x - cx, dy = y - cy, d2 = dx * dx + dy * dy;
      if (d2 > 1e-12) {
        if (d2 < best * best) best = std::sqrt(d2);
      } else {
        double m = jmin(jmin(jmin(x - rx, rx + rw - x), y - ry), ry + rh - y);
        if (-m < best) best = -m;
      }
    }
    if (!std::isfinite(best)) return false;
    d = best;
    return true;
  }
};

extern "C" kx_builder *kx_builder_new(void) { return new kx_builder(); }
extern "C" void kx_builder_free(kx_builder *b) { delete b; }

extern "C" int kx_builder_map(kx_builder *B, double width, double height, int32_t cols, int32_t rows, double density,
                              double T, uint32_t seed) {
  if (!B || cols < 1 || rows < 1 || !(width > 0) || !(height > 0)) return KX_EINVAL;
  B->map_w = width; B->map_h = height;
  auto &R = B->rects;
  R.clear();
  auto push = [&](double x, double y, double w, double h) { R.insert(R.end(), {x, y, w, h}); };
  push(0, 0, width, T);
  push(0, height - T, width, T);
  push(0, 0, T, height);
  push(width - T, 0, T, height);
  uint32_t s = seed;
  double cw = width / cols, ch = height / rows;
  for (int col = 1; col < cols; col++) {
    double xl = col * cw;
    for (int row = 0; row < rows; row++) {
      if (rng_next(s) > density) continue;
      double ys = row * ch;
      double gSz = jmax(60, jmin(ch * (0.28 + rng_next(s) * 0.38), height * 0.4));
      double gStart = ys + rng_next(s) * jmax(1, ch - gSz);
      double wt = 10 + std::floor(rng_next(s) * 10);
      if (gStart > ys + 14) push(xl - wt / 2, ys, wt, gStart - ys);
      double gEnd = gStart + gSz;
      if (gEnd < ys + ch - 14) push(xl - wt / 2, gEnd, wt, ys + ch - gEnd);
    }
  }
  for (int row = 1; row < rows; row++) {
    double yl = row * ch;
    for (int col = 0; col < cols; col++) {
      if (rng_next(s) > density) continue;
      double xs = col * cw;
      double gSz = jmax(60, jmin(cw * (0.28 + rng_next(s) * 0.38), width * 0.3));
      double gStart = xs + rng_next(s) * jmax(1, cw - gSz);
      double wt = 10 + std::floor(rng_next(s) * 10);
      if (gStart > xs + 14) push(xs, yl - wt / 2, gStart - xs, wt);
      double gEnd = gStart + gSz;
      if (gEnd < xs + cw - 14) push(gEnd, yl - wt / 2, xs + cw - gEnd, wt);
    }
  }
  return (int)(R.size() / 4);
}

extern "C" int kx_builder_find_spawn(kx_builder *B, double clearance, uint32_t *s, double *xy) {
  if (!B || !s || !xy) return KX_EINVAL;
  for (int i = 0; i < 500; i++) {
    double x = clearance + rng_next(*s) * jmax(1, B->map_w - clearance * 2);
    double y = clearance + rng_next(*s) * jmax(1, B->map_h - clearance * 2);
    double d;
    if (B->sdf(x, y, d) && d >= clearance) { xy[0] = x; xy[1] = y; return 1; }
  }
  xy[0] = B->map_w * 0.5; xy[1] = B->map_h * 0.5;
  return 0;
}

// crnSpawns (src/sim/train.js:103-122): training spawns on a ring around the
// walls. Each drawn slot tries up to 500 points in the map's 60 px inner
// margin for one whose fold distance d satisfies 26 <= d <= ARM_REACH (the
// mantle clear, a wall within the arms' reach), falls back to one unchecked
// point, then draws the heading; pair-shared slots 2p+1 copy slot 2p. With
// no map built the area is 240 x 240 and no point is tested.
extern "C" int kx_builder_crn_spawns(kx_builder *B, int32_t pop, int32_t pair_shared, uint32_t *s, double *out) {
  if (!B || !s || (pop > 0 && !out) || pop < 0) return KX_EINVAL;
  constexpr double kClear = 26, kReach = kShoulderR + 7 * kSegLen;
  const bool has_map = B->map_w > 0;
  const double W = has_map ? B->map_w : 240, H = has_map ? B->map_h : 240;
  int calls = 0;
  for (int i = 0; i < pop; i++) {
    double *o = out + 3 * (size_t)i;
    if (pair_shared && (i & 1)) { std::memcpy(o, o - 3, 3 * sizeof(double)); continue; }
    bool found = false;
    double x = 0, y = 0, d = 0;
    for (int t = 0; has_map && t < 500 && !found; t++, calls += 2) {
      x = 60 + rng_next(*s) * jmax(1, W - 120);
      y = 60 + rng_next(*s) * jmax(1, H - 120);
      found = B->sdf(x, y, d) && d >= kClear && d <= kReach;
    }
    if (!found) {
      x = 60 + rng_next(*s) * jmax(1, W - 120);
      y = 60 + rng_next(*s) * jmax(1, H - 120);
      calls += 2;
    }
    o[0] = x; o[1] = y; o[2] = rng_next(*s) * M_PI * 2;
    calls++;
  }
  return calls;
}

extern "C" int kx_builder_spawn(kx_builder *B, double x, double y, double angle, int32_t arms, int32_t segs,
                                int32_t ring, double curl) {
  if (!B || arms < 1 || arms > 8 || segs < 1 || segs > 15 || ring < 0) return KX_EINVAL;
  const double mantleI = (kMantleMass * kMantleR * kMantleR / 2) * kMantleIF;
  int mantle = B->add_body(x, y, angle, kMantleMass, mantleI);
  for (int i = 0; i < ring; i++) {
    double a = ((double)i / ring) * M_PI * 2;
    B->add_point(mantle, std::cos(a) * kMantleR * 0.7, std::sin(a) * kMantleR * 0.7, kMantleR * 0.45);
  }
  for (int a = 0; a < arms; a++) {
    double relAng = ((double)a / arms) * M_PI * 2;
    double th0 = angle + relAng;
    std::vector<int> sg(segs);
    double jx = x + std::cos(th0) * kShoulderR, jy = y + std::sin(th0) * kShoulderR, th = th0;
    for (int i = 0; i < segs; i++) {
      double t = segs > 1 ? (double)i / (segs - 1) : 0;
      double mass = kMassProx + (kMassDist - kMassProx) * t;
      double thick = kThickProx + (kThickDist - kThickProx) * t;
      double ux = std::cos(th), uy = std::sin(th);
      double cx = jx + ux * kSegLen * 0.5, cy = jy + uy * kSegLen * 0.5;
      int b = B->add_body(cx, cy, th, mass, mass * (kSegLen * kSegLen + thick * thick) / 12);
      double r = thick * 0.5;
      B->add_point(b, -kSegLen * 0.5 + r * 0.4, 0, r);
      B->add_point(b, kSegLen * 0.5 - r * 0.4, 0, r);
      sg[i] = b;
      jx = cx + ux * kSegLen * 0.5; jy = cy + uy * kSegLen * 0.5;
      th += curl;
    }
    B->add_cons(KX_PIN, mantle, sg[0], std::cos(relAng) * kShoulderR, std::sin(relAng) * kShoulderR, -kSegLen * 0.5, 0, 0);
    int ms = B->add_cons(KX_ANGLE, mantle, sg[0], 0, 0, 0, 0, 1 / kMuscleMin);
    B->rest[ms] = relAng; B->cmin[ms] = -kShoulderLimit; B->cmax[ms] = kShoulderLimit; B->cown[ms] = a;
    B->motors.push_back(ms);
    for (int i = 0; i < segs - 1; i++) {
      B->add_cons(KX_PIN, sg[i], sg[i + 1], kSegLen * 0.5, 0, -kSegLen * 0.5, 0, 0);
      int m = B->add_cons(KX_ANGLE, sg[i], sg[i + 1], 0, 0, 0, 0, 1 / kMuscleMin);
      B->rest[m] = curl; B->cmin[m] = -kJointLimit; B->cmax[m] = kJointLimit; B->cown[m] = a;
      B->motors.push_back(m);
    }
    int g = B->add_cons(KX_GRIP, sg[segs - 1], -1, kSegLen * 0.5, 0, x, y, 0);
    B->broken[g] = 1; B->bf[g] = 0; B->cown[g] = a;
    B->arm_tip.push_back(sg[segs - 1]);
    B->arm_grip.push_back(g);
    B->motor_off.push_back((int32_t)B->motors.size());
  }
  B->arm_off.push_back((int32_t)B->arm_tip.size());
  return (int)B->arm_off.size() - 2;
}

template <class T>
static void to_f32(std::vector<float> &dst, const std::vector<T> &src) {
  dst.resize(src.size());
  for (size_t i = 0; i < src.size(); i++) dst[i] = (float)src[i];
}

extern "C" int kx_builder_desc(kx_builder *B, kx_world_desc *d, kx_creature_desc *c) {
  if (!B || !d) return KX_EINVAL;
  to_f32(B->f_im, B->im); to_f32(B->f_ii, B->ii);
  to_f32(B->f_lax, B->lax); to_f32(B->f_lay, B->lay); to_f32(B->f_lbx, B->lbx); to_f32(B->f_lby, B->lby);
  to_f32(B->f_rest, B->rest); to_f32(B->f_min, B->cmin); to_f32(B->f_max, B->cmax);
  to_f32(B->f_plx, B->plx); to_f32(B->f_ply, B->ply); to_f32(B->f_pr, B->pr); to_f32(B->f_rects, B->rects);
  memset(d, 0, sizeof *d);
  d->n_bodies = (int32_t)B->px.size();
  d->n_constraints = (int32_t)B->ctype.size();
  d->n_points = (int32_t)B->pbody.size();
  d->n_rects = (int32_t)(B->rects.size() / 4);
  d->inv_mass = B->f_im.data(); d->inv_inertia = B->f_ii.data();
  d->c_type = B->ctype.data(); d->c_a = B->ca.data(); d->c_b = B->cb.data(); d->c_owner = B->cown.data();
  d->c_lax = B->f_lax.data(); d->c_lay = B->f_lay.data(); d->c_lbx = B->f_lbx.data(); d->c_lby = B->f_lby.data();
  d->c_rest = B->f_rest.data(); d->c_min = B->f_min.data(); d->c_max = B->f_max.data();
  d->c_color = nullptr;
  d->p_body = B->pbody.data(); d->p_lx = B->f_plx.data(); d->p_ly = B->f_ply.data(); d->p_r = B->f_pr.data();
  d->rects = B->f_rects.data();
  if (c) {
    c->n_creatures = (int32_t)B->arm_off.size() - 1;
    c->n_arms = (int32_t)B->arm_tip.size();
    c->creature_arm_off = B->arm_off.data();
    c->arm_tip = B->arm_tip.data(); c->arm_grip = B->arm_grip.data();
    c->arm_motor_off = B->motor_off.data(); c->motors = B->motors.data();
  }
  return 0;
}

extern "C" int kx_builder_state(const kx_builder *B, double *px, double *py, double *ang) {
  if (!B) return KX_EINVAL;
  size_t n = B->px.size();
  if (px) memcpy(px, B->px.data(), n * 8);
  if (py) memcpy(py, B->py.data(), n * 8);
  if (ang) memcpy(ang, B->ang.data(), n * 8);
  return 0;
}

extern "C" int kx_builder_cons_init(const kx_builder *B, double *lbx, double *lby, double *compliance, double *target,
                                    double *bfo, uint8_t *broken) {
  if (!B) return KX_EINVAL;
  size_t n = B->ctype.size();
  if (lbx) memcpy(lbx, B->lbx.data(), n * 8);
  if (lby) memcpy(lby, B->lby.data(), n * 8);
  if (compliance) memcpy(compliance, B->comp.data(), n * 8);
  if (target) memcpy(target, B->target.data(), n * 8);
  if (bfo) memcpy(bfo, B->bf.data(), n * 8);
  if (broken) memcpy(broken, B->broken.data(), n);
  return 0;
}
