// Morphology templates and their in-wave schedules for the persistent path.
//
// Creatures with identical static structure share one template (bodies'
// inverse mass/inertia, constraint topology and anchors, collision points,
// arms). Its constraints are scheduled into rounds for G lanes per creature:
// the colour order (src/physics/xpbd.js:108-126 + 290-296) is a sequential
// Gauss-Seidel order, and two constraints that share no body commute, so any
// topological order of the "shares a body, comes earlier" relation gives the
// same result bit for bit. Rounds are antichains of that relation of at most
// G constraints, all pins/grips or all angle motors (no divergent paths),
// filled by critical-path list scheduling.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/kxpbd.h"
#include "kx_sched.h"

// Pin and grip rounds hold at most G/2 constraints: each one is solved by a
// pair of adjacent lanes (one sin/cos per lane), see kx_pack_template.
int kx_pin_capacity(int G) { return G >= 2 ? G / 2 : 1; }

void kx_schedule_rounds(const TmplHost &t, int G, std::vector<int> &sched, std::vector<int> &rtype, int &R,
                        bool single) {
  const int nc = (int)t.type.size();
  std::vector<std::vector<int>> succ(nc);
  std::vector<int> npred(nc, 0), last(t.nb, -1);
  for (int k = 0; k < nc; k++) {
    int bodies[2] = {t.a[k], t.b[k]};
    int seen = -1;
    for (int j = 0; j < 2; j++) {
      int b = bodies[j];
      if (b < 0) continue;
      int p = last[b];
      if (p >= 0 && p != seen) { succ[p].push_back(k); npred[k]++; seen = p; }
      last[b] = k;
    }
  }
  // priority: longest chain to a sink
  std::vector<int> prio(nc, 1);
  for (int k = nc - 1; k >= 0; k--)
    for (int s : succ[k]) prio[k] = std::max(prio[k], prio[s] + 1);
  std::vector<int> left(npred);
  std::vector<char> done(nc, 0);
  std::vector<int> ready;
  for (int k = 0; k < nc; k++) if (!left[k]) ready.push_back(k);
  sched.clear(); rtype.clear();
  int scheduled = 0;
  while (scheduled < nc) {
    // pick the class (0: pin/grip, 1: angle) holding the most urgent ready node
    int best_cls = -1, best_p = -1, cnt_cls[2] = {0, 0};
    for (int k : ready) {
      int cls = t.type[k] == KX_ANGLE ? 1 : 0;
      cnt_cls[cls]++;
      if (prio[k] > best_p || (prio[k] == best_p && cls == 0)) { best_p = prio[k]; best_cls = cls; }
    }
    std::vector<int> cand;
    for (int k : ready) if ((t.type[k] == KX_ANGLE ? 1 : 0) == best_cls) cand.push_back(k);
    std::stable_sort(cand.begin(), cand.end(), [&](int x, int y) { return prio[x] > prio[y]; });
    const int cap = best_cls == 0 && !single ? kx_pin_capacity(G) : G;
    if ((int)cand.size() > cap) cand.resize(cap);
    std::vector<int> row(G, -1);
    for (int j = 0; j < (int)cand.size(); j++) row[j] = cand[j];
    sched.insert(sched.end(), row.begin(), row.end());
    rtype.push_back(best_cls);
    for (int k : cand) done[k] = 1;
    scheduled += (int)cand.size();
    std::vector<int> next;
    for (int k : ready) if (!done[k]) next.push_back(k);
    for (int k : cand)
      for (int s : succ[k]) if (--left[s] == 0) next.push_back(s);
    ready.swap(next);
  }
  R = (int)rtype.size();
}

void kx_schedule_collisions(const TmplHost &t, int G, std::vector<int> &lists, std::vector<int> &counts, int &maxbl) {
  std::vector<int> order(t.nb);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return t.pcnt[x] > t.pcnt[y]; });
  std::vector<std::vector<int>> lane(G);
  std::vector<int> load(G, 0);
  for (int b : order) {
    if (t.pcnt[b] == 0) continue;
    int best = 0;
    for (int l = 1; l < G; l++) if (load[l] < load[best]) best = l;
    lane[best].push_back(b);
    load[best] += t.pcnt[b];
  }
  maxbl = 1;
  for (auto &l : lane) maxbl = std::max<int>(maxbl, (int)l.size());
  lists.assign((size_t)G * maxbl, 0);
  counts.assign(G, 0);
  for (int l = 0; l < G; l++) {
    counts[l] = (int)lane[l].size();
    for (int j = 0; j < counts[l]; j++) lists[(size_t)l * maxbl + j] = lane[l][j];
  }
}

static void put(std::vector<uint32_t> &pool, const std::vector<int> &v) {
  for (int x : v) pool.push_back((uint32_t)x);
}
static uint32_t fbits(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  return u;
}

// Serialise one template with its schedule for G lanes into the word pool.
// Sections start on 16-byte boundaries; constraint, body and point records
// are 16-byte vectors so the kernel reads each with one LDS instruction:
//   crec[k] = {lax, lay, -, cab} (pin, grip) or {rest, min, max, cab}
//             (angle), cab = a | b << 8 | type << 16            (16 B)
//   brec[b] = {im, ii, point offset | count << 16, contact radius}
//   prec[q] = {lx, ly, r, body}
KxTmpl kx_pack_template(const TmplHost &t, int G, std::vector<uint32_t> &pool, bool single) {
  KxTmpl T{};
  T.nb = t.nb; T.nc = (int)t.type.size(); T.np = (int)t.plx.size();
  T.narms = (int)t.arm_tip.size(); T.nmot = (int)t.mot_k.size(); T.ngrip = (int)t.grip_k.size();
  std::vector<int> sched, rtype, lists, counts;
  kx_schedule_rounds(t, G, sched, rtype, T.R, single);
  T.single = single ? 1 : 0;
  kx_schedule_collisions(t, G, lists, counts, T.maxbl);
  while (pool.size() % 4) pool.push_back(0u);
  T.base = (int)pool.size();
  auto at = [&]() {
    while ((pool.size() - T.base) % 4) pool.push_back(0u);
    return (int)pool.size() - T.base;
  };
  T.o_crec = at();
  for (int k = 0; k < T.nc; k++) {
    int cab = t.a[k] | ((t.b[k] < 0 ? 255 : t.b[k]) << 8) | (t.type[k] << 16);
    if (t.type[k] == KX_ANGLE)
      pool.insert(pool.end(), {fbits(t.rest[k]), fbits(t.mn[k]), fbits(t.mx[k]), (uint32_t)cab});
    else
      pool.insert(pool.end(), {fbits(t.lax[k]), fbits(t.lay[k]), 0u, (uint32_t)cab});
  }
  T.o_brec = at();
  for (int b = 0; b < t.nb; b++) {
    // contact radius: the largest distance from the body's centre at which
    // one of its collision points can touch a rect (offset length + radius),
    // plus 0.05 px for the binary32 rounding of world points and distances,
    // rounded up to binary32
    double reach = 0.0;
    for (int j = 0; j < t.pcnt[b]; j++) {
      int q = t.poff[b] + j;
      reach = std::max(reach, std::sqrt((double)t.plx[q] * t.plx[q] + (double)t.ply[q] * t.ply[q]) +
                                  std::fabs((double)t.pr[q]));
    }
    const double want = reach + 0.05;
    float cr = (float)want;
    if ((double)cr < want) cr = std::nextafter(cr, INFINITY);
    if (!std::isfinite(reach)) cr = INFINITY;
    pool.insert(pool.end(), {fbits(t.im[b]), fbits(t.ii[b]), (uint32_t)(t.poff[b] | (t.pcnt[b] << 16)), fbits(cr)});
  }
  T.o_prec = at();
  std::vector<int> pbody(t.plx.size(), 0);
  for (int b = 0; b < t.nb; b++)
    for (int j = 0; j < t.pcnt[b]; j++) pbody[t.poff[b] + j] = b;
  for (size_t q = 0; q < t.plx.size(); q++)
    pool.insert(pool.end(), {fbits(t.plx[q]), fbits(t.ply[q]), fbits(t.pr[q]), (uint32_t)pbody[q]});
  // round slots as packed descriptors: k | A << 8 | B << 16 (B = 255: world
  // anchor), -1 for an idle lane, so one LDS read yields every index; a pin
  // round's j-th constraint occupies lanes 2j and 2j+1 (lane j when single).
  // Stored by groups of four rounds, lane-major within a group:
  // desc4[(grp * G + lane) * 4 + r % 4], so a lane reads a group's four
  // slots with one 16-byte load.
  // A motor whose two bodies are those of a pin in the round before (a
  // joint's hinge and its muscle) sits on that pin's even lane (its lane when
  // single), the others on the lowest free lanes; lanes within a round are
  // interchangeable, and this placement lets a kernel solve the motor right
  // after the pin on the pose it holds (the fused rounds, kx_persist.hip).
  const int R4 = (T.R + 3) / 4;
  std::vector<int> desc((size_t)R4 * G * 4, -1);
  auto put_slot = [&](int r, int lane, int v) { desc[((size_t)(r >> 2) * G + lane) * 4 + (r & 3)] = v; };
  auto slot_of = [&](int r, int lane) { return desc[((size_t)(r >> 2) * G + lane) * 4 + (r & 3)]; };
  auto pack = [&](int k) { return k | (t.a[k] << 8) | ((t.b[k] < 0 ? 255 : t.b[k]) << 16); };
  for (int r = 0; r < T.R; r++) {
    if (!rtype[r]) {
      for (int j = 0; j < G; j++) {
        const int k = sched[(size_t)r * G + j];
        if (k < 0) continue;
        if (single) { put_slot(r, j, pack(k)); continue; }
        put_slot(r, 2 * j, pack(k)); put_slot(r, 2 * j + 1, pack(k));
      }
      continue;
    }
    std::vector<char> used(G, 0);
    std::vector<int> rest;
    for (int j = 0; j < G; j++) {
      const int k = sched[(size_t)r * G + j];
      if (k < 0) continue;
      int lane = -1;
      if (r > 0 && !rtype[r - 1] && t.b[k] >= 0)
        for (int l = 0; l < G; l += single ? 1 : 2) {
          const int p = slot_of(r - 1, l);
          if (p >= 0 && !used[l] && (p >> 8) == (pack(k) >> 8)) { lane = l; break; }
        }
      if (lane >= 0) { put_slot(r, lane, pack(k)); used[lane] = 1; }
      else rest.push_back(k);
    }
    for (int k : rest)
      for (int l = 0; l < G; l++)
        if (!used[l]) { put_slot(r, l, pack(k)); used[l] = 1; break; }
  }
  T.o_sched = at(); put(pool, desc);
  // round types as a bitmask, 32 rounds per word (bit set = angle round)
  std::vector<int> rbits(std::max<size_t>((rtype.size() + 31) / 32, 2), 0);   // at least two words
  for (size_t r = 0; r < rtype.size(); r++)
    if (rtype[r]) rbits[r >> 5] |= (int)(1u << (r & 31));
  T.o_rtype = at(); put(pool, rbits);
  // (the per-lane collision lists only set maxbl, the workgroup cost order;
  // the kernel finds its collision candidates itself)
  T.o_coll = T.o_ccnt = 0;
  T.o_arm_tip = at(); put(pool, t.arm_tip);
  T.o_arm_grip = at(); put(pool, t.arm_grip);
  T.o_mot_k = at(); put(pool, t.mot_k);
  T.o_mot_slot = at(); put(pool, t.mot_slot);
  T.o_grip_k = at(); put(pool, t.grip_k);
  T.o_grip_arm = at(); put(pool, t.grip_arm);
  std::vector<int> arm_gslot(t.arm_grip.size(), -1);
  for (size_t a = 0; a < t.arm_grip.size(); a++)
    for (size_t j = 0; j < t.grip_k.size(); j++)
      if (t.grip_k[j] == t.arm_grip[a]) arm_gslot[a] = (int)j;
  T.o_arm_gslot = at(); put(pool, arm_gslot);
  T.words = at();
  return T;
}

std::string kx_template_key(const TmplHost &t) {
  std::string s;
  auto add = [&](const void *p, size_t n) { s.append((const char *)p, n); };
  auto addv = [&](const auto &v) {
    size_t n = v.size();
    add(&n, sizeof n);
    if (n) add(v.data(), n * sizeof(v[0]));
  };
  add(&t.nb, sizeof t.nb);
  addv(t.type); addv(t.a); addv(t.b); addv(t.lax); addv(t.lay); addv(t.rest); addv(t.mn); addv(t.mx);
  addv(t.im); addv(t.ii); addv(t.poff); addv(t.pcnt); addv(t.plx); addv(t.ply); addv(t.pr);
  addv(t.arm_tip); addv(t.arm_grip); addv(t.mot_k); addv(t.mot_slot); addv(t.grip_k); addv(t.grip_arm);
  return s;
}
