// Persistent batched path. One workgroup holds WPB wavefronts; each
// wavefront steps 64/G creatures of one morphology template, G lanes per
// creature, with every creature's bodies and constraint state in LDS for all
// substeps and iterations of a multi-step launch. The template (constraint
// topology, anchors, collision points, the round schedule) is copied to LDS
// once per workgroup and shared by its waves.
//
// A colour pass of the reference (one sequential dispatch per colour,
// src/gpu/gpuworld.js:353-366) becomes a list of in-wave rounds built on the
// host (kx_sched.cpp): a wave's LDS traffic is processed in issue order, so a
// compiler fence between rounds is all the ordering they need. State is
// read from and written to HBM (SoA) only at launch boundaries.
//
// Within a step, positions are relative to an integer origin (rint of the
// creature's first body, taken after the control law); the subtraction is
// exact for coordinates of magnitude >= 1, so the solver works with finer
// position ulps. Between steps the state is global binary32, exactly as the
// plain path (kx_plain.hip) holds it, and the two paths agree bit for bit.
#include <atomic>

#include "../../include/kxpbd.h"
#include "kx_device.h"
#include "kx_map_lds.h"
#include "kx_sched.h"

#define KX_WAVE_SYNC()                                     \
  do {                                                     \
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
    __builtin_amdgcn_wave_barrier();                       \
  } while (0)

// Forces every component of a just-loaded 16-byte record into registers at
// this point, so the record is read by one full-width LDS load issued with
// its neighbours rather than narrowed and re-issued inside later branches.
typedef float kx_f4v __attribute__((ext_vector_type(4)));
typedef float kx_f3v __attribute__((ext_vector_type(3)));   // 16-byte aligned: one 96-bit LDS access
typedef float kx_f2v __attribute__((ext_vector_type(2)));
#define KX_HOLD4(v)                          \
  do {                                       \
    kx_f4v t_ = {(v).x, (v).y, (v).z, (v).w}; \
    asm volatile("" : "+v"(t_));             \
    (v) = make_float4(t_.x, t_.y, t_.z, t_.w); \
  } while (0)

// A round slot's records other than the body poses: dyn, lambda, the
// constraint record (3 words: both round types' fields) and both bodies'
// {im, ii}, with the decoded slot. A lane with no slot (dsc < 0) reads
// constraint 0 and body 0 (every read stays unconditional) and is not live.
// bodies per lane whose broad-phase reads are batched (sub + k*G, k < KX_NBL)
#define KX_NBL 4

struct KxSlot {
  float4 dk;
  kx_f2v lk;
  kx_f3v cr;
  kx_f2v bA, bB;
  int k, A, B;
  bool hasB, live;
};

__device__ __forceinline__ KxSlot kx_slot_read(int dsc, const float4 *dyn, const kx_f2v *lam2, const float4 *t_crec,
                                               const float4 *t_brec) {
  KxSlot s;
  s.live = dsc >= 0;
  const int dl = s.live ? dsc : 0;
  s.k = dl & 255;
  s.A = (dl >> 8) & 255;
  const int B0 = (dl >> 16) & 255;
  s.hasB = B0 != 255;
  s.B = s.hasB ? B0 : s.A;
  s.dk = dyn[s.k];
  s.lk = lam2[s.k];
  s.cr = *(const kx_f3v *)&t_crec[s.k];
  s.bA = *(const kx_f2v *)&t_brec[s.A];
  s.bB = *(const kx_f2v *)&t_brec[s.B];
  return s;
}

// per-creature LDS layout (word offsets; records of 16 or 8 bytes):
// Each body's start-of-substep pose and velocity stay in the registers of the
// lane that owns the body (lane i % G, at most KX_BODY_REGS bodies per lane);
// LDS holds what other lanes read:
//   pos[i] = {ppx, ppy, angP, -}   (predicted pose, read by every solve)
//   tip[a] = {px, py, ang, vx}, {vy, -, -, -}   (arm a's tip body, for the reflex)
//   dyn[k] = {alpha, broken, lbx, lby} (pin, grip) or {alpha, target, broken, 1/(wA+wB)} (angle):
//            each type keeps its broken flag in the word it does not use
//   lam[k] = {lambda_x, lambda_y} (pin, grip) or {lambda, -} (angle)
//   gbf[j] = break force of the j-th grip (the only constraints that break)
//   ev     = {step events, collision candidate count, -, -}
// and, for launches that fold collisions in parallel (P.coll_par):
//   flg    = one bit per collision point: "corrects at the pose after the rounds"
//   cand   = collision candidates: point | body << 16
struct KxCrLds {
  int pos, tip, dyn, lam, gbf, ev, flg, cand, words;
  __host__ __device__ KxCrLds(int NB, int NC, int NP, int NG, int NA, bool par) {
    int o = 0;
    pos = o; o += 4 * NB; tip = o; o += 8 * NA;
    dyn = o; o += 4 * NC; lam = o; o += (2 * NC + 3) / 4 * 4;
    gbf = o; o += (NG + 3) / 4 * 4;
    ev = o; o += 4;
    flg = o; o += par ? ((NP + 31) / 32 + 3) / 4 * 4 : 0;
    cand = o; o += par ? (NP + 3) / 4 * 4 : 0;
    words = o;
  }
};

int kx_persist_cr_words(int NB, int NC, int NP, int NG, int NA, bool par) {
  return KxCrLds(NB, NC, NP, NG, NA, par).words;
}

// Whether constraint record `v` (crec.w) is a world-anchored pin or grip.
__device__ __forceinline__ bool kx_world_anchored(int v) { return ((v >> 8) & 255) == 255 && (v >> 16) != KX_ANGLE; }

// Adds (sx, sy) to the world anchors of a creature's world-anchored pins and
// grips; with the same additions to the body poses (kx_persist_body) these
// are the plain path's shift in binary32 (an angle constraint has no anchor).
// wa: the lane's world-anchored constraints among sub + j*G (bit j), or ~0u
// when the creature has more than 32 constraints per lane (then every
// constraint of the lane is tested).
__device__ __forceinline__ void kx_shift_anchors(float4 *dyn, const float4 *t_crec, int nc, int sub, int G, uint32_t wa,
                                                 float sx, float sy) {
  if (wa == ~0u) {
    for (int k = sub; k < nc; k += G) {
      if (kx_world_anchored(__float_as_int(t_crec[k].w))) {
        float4 dk = dyn[k];
        dk.z += sx; dk.w += sy;
        dyn[k] = dk;
      }
    }
    return;
  }
  for (uint32_t m = wa; m; m &= m - 1u) {
    const int k = sub + __builtin_ctz(m) * G;
    float4 dk = dyn[k];
    dk.z += sx; dk.w += sy;
    dyn[k] = dk;
  }
}

// Value of the even (kx_pair_even) or odd (kx_pair_odd) lane of an adjacent
// pair (lanes 2j, 2j+1) on both lanes, one DPP move each (quad_perm
// [0,0,2,2] and [1,1,3,3]).
__device__ __forceinline__ float kx_pair_even(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xA0, 0xF, 0xF, true));
}
__device__ __forceinline__ float kx_pair_odd(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xF5, 0xF, 0xF, true));
}

// Register-held round slot (RR > 0): this lane's slot of one round, its
// records read from LDS once per launch and its dynamics once per step.
struct KxRegSlot {
  float4 dk;
  kx_f2v lk;
  kx_f3v cr;
  kx_f2v bA, bB;
  int A, B;
  float rm;   // motor: 1 / (wA + wB + alpha), refreshed with alpha once per step
};

// The body of one workgroup `bid`: WPB wavefronts of 64/G creatures of one
// template (workgroup tables P.wg_*). RR > 0: every template of the launch
// has the fixed schedule of RR rounds (KX_REG_RTYPE, fused hinge and muscle
// rounds); with REG each lane keeps its RR round slots in registers
// (records, dynamics, lambda), so a round reads only its two poses from LDS,
// without it a fused round reads its two slots' records from LDS.
template <int G, int WPB, bool PAR, int RR = 0, bool REG = true>
__device__ __forceinline__ void kx_persist_body(const KxDev &d, const KxLaw &L, int has_law, int step0,
                                                const int *step_base, int nsteps, int ev_row0, int with_forces,
                                                const KxPersistCfg &P, int bid) {
  constexpr int CPW = KX_WAVE / G;
  extern __shared__ __attribute__((aligned(16))) uint32_t smem[];
  const int lane = threadIdx.x & (KX_WAVE - 1);
  const int q = lane / G, sub = lane % G;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / KX_WAVE);

  // ── template + cell bitmap -> LDS (the only workgroup barrier) ──
  const int tid = P.wg_tmpl[bid];
  const KxTmpl T = P.tmpl[tid];
  for (int i = threadIdx.x; i < T.words; i += blockDim.x) smem[i] = P.pool[T.base + i];
  uint32_t *bits = smem + P.tmpl_words;
  for (int i = threadIdx.x; i < P.bm_words; i += blockDim.x) bits[i] = d.cell_bits[i];
  // the map when it fits: the rect of every cell entry, then cell offsets
  float4 *m_crect = (float4 *)(smem + P.tmpl_words + P.bm_words);
  uint16_t *m_off = (uint16_t *)(smem + P.tmpl_words + P.bm_words + 4 * P.map_ids);
  if (P.map_words) {
    for (int i = threadIdx.x; i < P.map_ids; i += blockDim.x) m_crect[i] = d.cell_rects[i];
    for (int i = threadIdx.x; i <= P.map_cells; i += blockDim.x) m_off[i] = (uint16_t)d.cell_off[i];
  }
  __syncthreads();

  const float4 *t_crec = (const float4 *)(smem + T.o_crec);   // one per constraint
  const float4 *t_brec = (const float4 *)(smem + T.o_brec);
  const float4 *t_prec = (const float4 *)(smem + T.o_prec);
  const int4 *t_sched4 = (const int4 *)(smem + T.o_sched);
  const int *t_rbits = (const int *)(smem + T.o_rtype);
  const int *t_arm_tip = (const int *)(smem + T.o_arm_tip), *t_arm_grip = (const int *)(smem + T.o_arm_grip);
  const int *t_mot_k = (const int *)(smem + T.o_mot_k), *t_mot_slot = (const int *)(smem + T.o_mot_slot);
  const int *t_grip_k = (const int *)(smem + T.o_grip_k), *t_grip_arm = (const int *)(smem + T.o_grip_arm);
  const int *t_arm_gslot = (const int *)(smem + T.o_arm_gslot);

  const int cid = P.wave_cr[P.wg_cr_off[bid] + w * CPW + q];
  const bool act = cid >= 0;
  const KxCrLds Lo(T.nb, T.nc, T.np, T.ngrip, T.narms, PAR);
  uint32_t *S = smem + P.tmpl_words + P.bm_words + P.map_words + (size_t)(w * CPW + q) * T.cr_stride;
  float4 *pos = (float4 *)(S + Lo.pos), *tip = (float4 *)(S + Lo.tip);
  float4 *dyn = (float4 *)(S + Lo.dyn);
  kx_f2v *lam2 = (kx_f2v *)(S + Lo.lam);
  float *gbf = (float *)(S + Lo.gbf);
  uint32_t *evw = S + Lo.ev, *flg = S + Lo.flg;
  uint32_t *cand = S + Lo.cand;
  const int nb = act ? T.nb : 0, nc = act ? T.nc : 0;
  const int b0 = act ? d.cr[cid].body_off : 0, c0 = act ? d.cr[cid].cons_off : 0;

  // ── load the creature's dynamic state (global coordinates) ──
  // pose (rx, ry, ra) and velocity (rvx, rvy, rva) of the lane's bodies
  // sub + k*G, k < BR, in registers; every loop over them is unrolled (the
  // host guarantees nb <= BR * G: KX_BODY_REGS, or KX_REG_BODY_REGS for the
  // register-slot kernel, whose slots leave less of the register file)
  constexpr int BR = RR > 0 ? KX_REG_BODY_REGS : KX_BODY_REGS;
  float rx[BR], ry[BR], ra[BR], rvx[BR], rvy[BR], rva[BR];
  // bodies per lane of this workgroup's template (wave-uniform): register
  // slots past it are skipped by a scalar branch
  const int nbl = __builtin_amdgcn_readfirstlane((T.nb + G - 1) / G);
#pragma unroll
  for (int k = 0; k < BR; k++) {
    const int i = sub + k * G;
    rx[k] = ry[k] = ra[k] = rvx[k] = rvy[k] = rva[k] = 0.0f;
    if (i < nb) {
      const int g = b0 + i;
      rx[k] = d.px[g]; ry[k] = d.py[g]; ra[k] = d.ang[g];
      rvx[k] = d.vx[g]; rvy[k] = d.vy[g]; rva[k] = d.av[g];
    }
  }
  // each arm's tip body, published by its owning lane for the reflex lane
  auto publish_tips = [&]() {
    for (int a = 0; a < (act ? T.narms : 0); a++) {
      const int t = t_arm_tip[a];
      if (t % G != sub) continue;
      const int kt = t / G;
      float x = 0.0f, y = 0.0f, an = 0.0f, vx = 0.0f, vy = 0.0f;
#pragma unroll
      for (int k = 0; k < BR; k++) {
        if (k >= nbl) break;
        if (k == kt) { x = rx[k]; y = ry[k]; an = ra[k]; vx = rvx[k]; vy = rvy[k]; }
      }
      tip[2 * a] = make_float4(x, y, an, vx);
      tip[2 * a + 1] = make_float4(vy, 0.0f, 0.0f, 0.0f);
    }
  };
  publish_tips();
  for (int k = sub; k < nc; k += G) {
    const int g = c0 + k;
    const float alpha = d.comp[g] * d.invH2, brk = __int_as_float(d.broken[g]);
    const int cab = __float_as_int(t_crec[k].w), kA = cab & 255, kB = (cab >> 8) & 255;
    // an angle's 1 / (wA + wB), as kx_angle_core would compute it
    const float rS = kx_rcp(t_brec[kA].y + (kB != 255 ? t_brec[kB].y : 0.0f));
    dyn[k] = (cab >> 16) == KX_ANGLE ? make_float4(alpha, d.target[g], brk, rS)
                                     : make_float4(alpha, brk, d.lbx[g], d.lby[g]);
    lam2[k] = kx_f2v{0.0f, 0.0f};
  }
  for (int j = sub; j < (act ? T.ngrip : 0); j += G) gbf[j] = d.bf[c0 + t_grip_k[j]];
  if constexpr (PAR) {
    // collision flags start at zero; each pass that sets some clears them
    for (int wi = sub; wi < (act ? (T.np + 31) >> 5 : 0); wi += G) flg[wi] = 0u;
  }
  double my_cd = 0.0;
  if (act && sub < T.narms) my_cd = d.cd[cid * KX_MAX_ARMS + sub];
  const double hS = d.hd * d.S;
  // the lane's collidable bodies among sub + k*G, k < NBL (dynamic, with
  // collision points): bit k; bodies past NBL*G are tested one by one (the
  // register-slot kernel batches all of its at most BR bodies per lane)
  constexpr int NBL = RR > 0 ? BR : KX_NBL;
  uint32_t bp_mask = 0u;
#pragma unroll
  for (int k = 0; k < NBL; k++) {
    const int i = sub + k * G;
    if (act && i < nb) {
      const float4 br = t_brec[i];
      if (br.x != 0.0f && (__float_as_int(br.z) >> 16) > 0) bp_mask |= 1u << k;
    }
  }
  // the lane's world-anchored pins and grips among sub + j*G (bit j)
  uint32_t wa_mask = 0u;
  if (nc > 32 * G) wa_mask = ~0u;
  else
    for (int j = 0, k = sub; k < nc; j++, k += G)
      if (kx_world_anchored(__float_as_int(t_crec[k].w))) wa_mask |= 1u << j;
  double comp_last = L.comp;      // motor compliance of the latest control step
  // round types of the first 64 rounds (bit r set = angle round)
  const uint32_t rb0 = (uint32_t)__builtin_amdgcn_readfirstlane(t_rbits[0]);
  const uint32_t rb1 = (uint32_t)__builtin_amdgcn_readfirstlane(t_rbits[1]);
  KX_WAVE_SYNC();

  // register-held round slots: the slot descriptor of round r for this lane
  // (-1 idle), and the lane's records; bit r of rflags: live and not broken
  // (set per step), bit 16 + r: a second body
  constexpr int NR = RR > 0 ? RR : 1;
  [[maybe_unused]] KxRegSlot slot[NR];
  [[maybe_unused]] uint32_t rflags = 0u;
  auto slot_dsc = [&](int r) {
    const int4 d4 = act ? t_sched4[(r >> 2) * G + sub] : make_int4(-1, -1, -1, -1);
    return (r & 3) == 0 ? d4.x : (r & 3) == 1 ? d4.y : (r & 3) == 2 ? d4.z : d4.w;
  };
  // fixed schedule with LDS-read slots: every round's descriptor, held
  [[maybe_unused]] int fdsc[NR];
  if constexpr (RR > 0 && !REG) {
#pragma unroll
    for (int r = 0; r < RR; r++) fdsc[r] = slot_dsc(r);
  }
  // One fused round of LDS-held slots: a pin (or grip) on lane pairs and,
  // on the even lane, the motor with the same two bodies solved on the pose
  // the pin leaves (its own pose if the pin does not commit), which is the
  // pose the motor's round would read; the even lane writes the pair's two
  // bodies once. pinB: every pin of the round has a second body (a
  // constant of the fixed schedule). single (a template packed with each pin
  // on one lane): every lane takes both sin/cos itself and commits its own
  // pin and motor; the values are those of the pair form.
  [[maybe_unused]] auto fused_pair = [&](int dc, int dm, bool pinB, bool single) {
    const KxSlot c = kx_slot_read(dc, dyn, lam2, t_crec, t_brec);
    const KxSlot m = kx_slot_read(dm, dyn, lam2, t_crec, t_brec);
    const kx_f3v pA = *(const kx_f3v *)&pos[c.A], pB = *(const kx_f3v *)&pos[c.B];
    // pins and grips keep their broken flag in dyn.y, motors in dyn.z
    const bool livep = c.live && __float_as_int(c.dk.y) == 0;
    const bool livem = m.live && __float_as_int(m.dk.z) == 0;
    const bool hasB = pinB || c.hasB;
    const bool odd = !single && (sub & 1) != 0;
    float sa, ca, sb, cb;
    if (single) {
      kx_sincos(pA.z, &sa, &ca);
      kx_sincos(pB.z, &sb, &cb);
    } else {
      float s1, c1;
      kx_sincos(odd ? pB.z : pA.z, &s1, &c1);
      sa = kx_pair_even(s1); ca = kx_pair_even(c1); sb = kx_pair_odd(s1); cb = kx_pair_odd(c1);
    }
    float xA = pA.x, yA = pA.y, tA = pA.z, xB = pB.x, yB = pB.y, tB = pB.z;
    float lx = c.lk.x, ly = c.lk.y;
    const bool okp = kx_pin_core_sc(xA, yA, tA, c.bA.x, c.bA.y, xB, yB, tB, c.bB.x, c.bB.y, hasB, c.cr.x, c.cr.y,
                                    c.dk.z, c.dk.w, c.dk.x, lx, ly, sa, ca, sb, cb) &&
                     livep;
    xA = okp ? xA : pA.x; yA = okp ? yA : pA.y; tA = okp ? tA : pA.z;
    xB = okp ? xB : pB.x; yB = okp ? yB : pB.y; tB = okp ? tB : pB.z;
    float uA = tA, uB = tB, lam = m.lk.x;
    const float rm = kx_rcp((m.bA.y + m.bB.y) + m.dk.x);
    const bool okm =
        kx_angle_core_ab(uA, uB, m.bA.y, m.bB.y, m.cr.x, m.dk.y, m.cr.y, m.cr.z, m.dk.x, m.dk.w, rm, lam) && livem;
    if (okm) { tA = uA; tB = uB; ((float *)&lam2[m.k])[0] = lam; }
    if (!odd && okp) lam2[c.k] = kx_f2v{lx, ly};
    if (!odd && (okp || okm)) {
      // B first: for a world anchor B == A and pA must win
      *(kx_f3v *)&pos[c.B] = kx_f3v{xB, yB, tB};
      *(kx_f3v *)&pos[c.A] = kx_f3v{xA, yA, tA};
    }
    KX_WAVE_SYNC();
  };
  if constexpr (RR > 0 && REG) {
#pragma unroll
    for (int r = 0; r < RR; r++) {
      const KxSlot c = kx_slot_read(slot_dsc(r), dyn, lam2, t_crec, t_brec);
      slot[r] = KxRegSlot{c.dk, c.lk, c.cr, c.bA, c.bB, c.A, c.B, 0.0f};
      rflags |= (c.hasB ? 1u : 0u) << (16 + r);
    }
  }

  for (int step = 0; step < nsteps; step++) {
    const int key = (step_base ? step_base[0] : step0) + step;
    if (act && sub == 0) evw[0] = 0u;
    uint32_t my_ev = 0u;
    if (has_law) {
      // motor targets + compliance (core.js:190-198), from the seeded law or
      // from this step's tape record
      const bool tape = (L.flags & KX_LAW_TAPE) != 0;
      const float *rec = tape && act ? kx_tape_rec(d, step, cid) : nullptr;
      if (rec) comp_last = kx_tape_comp(L, rec[0]);
      const float a_law = (float)comp_last * d.invH2;
      const uint32_t ck = (uint32_t)cid + L.cbase;
      const uint32_t hstep = kx_ctl_prefix(L.seed, ck, (uint32_t)key);   // this creature's step in the chain
      for (int j = sub; j < (act ? T.nmot : 0); j += G) {
        int k = t_mot_k[j];
        float4 dk = dyn[k];
        dk.x = a_law;
        dk.y = rec ? kx_tape_target(L, rec, t_mot_slot[j]) : kx_law_target_at(L, hstep, t_mot_slot[j]);
        dyn[k] = dk;
      }
      // adhesion reflex (core.js:199-218), one lane per arm
      if (act && sub < T.narms) {
        double adh = rec ? (double)rec[3 + 3 * sub] : kx_ctl_u_at(hstep, KX_SLOT_ADH + sub);
        double gu = (L.flags & KX_LAW_FIXED_GRIP) ? kx_ctl_u(L.seed, ck, KX_GRIP_DRAW_KEY, KX_SLOT_ADH + sub) : 0.0;
        const int g = t_arm_grip[sub], j = t_arm_gslot[sub];
        float4 dk = dyn[g], tp = tip[2 * sub], cr0 = t_crec[g];   // tip {px, py, ang, vx}; crec {lax, lay, -, cab}
        const float4 tv = make_float4(tp.w, tip[2 * sub + 1].x, 0.0f, 0.0f);   // tip {vx, vy}
        const kx_f2v lk = lam2[g];
        int bk = __float_as_int(dk.y);
        float f_bf = gbf[j], gx = dk.z, gy = dk.w, lx = lk.x, ly = lk.y;
        uint32_t e = kx_reflex(L, hS, adh, gu, my_cd, bk, f_bf, gx, gy, lx, ly, tv.x, tv.y, (double)tp.x, (double)tp.y,
                               tp.z, cr0.x, cr0.y);
        dyn[g] = make_float4(dk.x, __int_as_float(bk), gx, gy);
        lam2[g] = kx_f2v{lx, ly};
        gbf[j] = f_bf;
        my_ev |= e << (3 * sub);
      }
      KX_WAVE_SYNC();
    }
    // Between steps the state is held in global binary32 coordinates, as the
    // reference's buffers hold it; a step works relative to the integer
    // origin rint(first body) (0 if that is not finite), subtracted before
    // and added back after.
    // (body 0 is lane 0's first body: its origin reaches the other lanes
    // through two spare event words)
    if (act && sub == 0) { evw[2] = __float_as_uint(kx_origin(rx[0])); evw[3] = __float_as_uint(kx_origin(ry[0])); }
    KX_WAVE_SYNC();
    const float ox = act ? __uint_as_float(evw[2]) : 0.0f, oy = act ? __uint_as_float(evw[3]) : 0.0f;
#pragma unroll
    for (int k = 0; k < BR; k++) {
      if (k >= nbl) break;
      rx[k] -= ox; ry[k] -= oy;
    }
    kx_shift_anchors(dyn, t_crec, nc, sub, G, wa_mask, -ox, -oy);
    KX_WAVE_SYNC();
    if constexpr (RR > 0 && REG) {
      // the step's dynamics (control law, reflex, origin shift) into the
      // slots, and which slots commit this step: live and not broken (an
      // angle keeps its broken flag in dyn.z, a pin or grip in dyn.y), and
      // for a motor wA + wB > 0, the test its solve would otherwise repeat
      uint32_t lv = 0u;
#pragma unroll
      for (int r = 0; r < RR; r++) {
        const int dsc = slot_dsc(r);
        const float4 dk = dyn[dsc >= 0 ? dsc & 255 : 0];
        slot[r].dk = dk;
        const bool is_ang = ((KX_REG_RTYPE >> r) & 1) != 0;
        if (is_ang) slot[r].rm = kx_rcp((slot[r].bA.y + slot[r].bB.y) + dk.x);
        const int brk = __float_as_int(is_ang ? dk.z : dk.y);
        const bool wpos = !is_ang || (slot[r].bA.y + slot[r].bB.y) > 0.0f;
        lv |= (dsc >= 0 && brk == 0 && wpos ? 1u : 0u) << r;
      }
      rflags = (rflags & 0xFFFF0000u) | lv;
    }
    for (int s = 0; s < d.S; s++) {
      // resetLambda + integrate (xpbd.js:262-287)
      if constexpr (RR > 0 && REG) {
#pragma unroll
        for (int r = 0; r < RR; r++) slot[r].lk = kx_f2v{0.0f, 0.0f};
      } else {
        for (int k = sub; k < nc; k += G) lam2[k] = kx_f2v{0.0f, 0.0f};
      }
      const bool forces = with_forces && step == 0 && s == 0;
#pragma unroll
      for (int k = 0; k < BR; k++) {
        if (k >= nbl) break;
        const int i = sub + k * G;
        if (i >= nb) continue;
        float fx = 0.0f, fy = 0.0f, tq = 0.0f;
        if (forces) {
          int g = b0 + i;
          fx = d.fx[g]; fy = d.fy[g]; tq = d.tq[g];
          d.fx[g] = 0.0f; d.fy[g] = 0.0f; d.tq[g] = 0.0f;
        }
        const float4 br = t_brec[i];
        float q_x, q_y, q_a;
        kx_integrate(d, rvx[k], rvy[k], rva[k], rx[k], ry[k], ra[k], br.x, br.y, fx, fy, tq, q_x, q_y, q_a);
        pos[i] = make_float4(q_x, q_y, q_a, 0.0f);
      }
      KX_WAVE_SYNC();
      for (int it = 0; it < d.I; it++) {
        // constraint rounds (type-homogeneous, wave-uniform branch), four per
        // group: a lane reads the group's four slot descriptors with one load,
        // and the next group's descriptors are fetched while this group runs.
        // The rounds are software-pipelined: a round first reads its two
        // body poses (which the previous round may have written), then
        // issues the next round's other records (dyn, lambda, constraint and
        // body records: none of them is written by a round), and waits only
        // for the poses. A broken constraint is solved too and its result
        // dropped.
        if constexpr (RR > 0 && REG) {
          // the same solves on the register-held slots, a joint's hinge and
          // its muscle fused: rounds 2f (pins and grips on lane pairs) and
          // 2f + 1 (motors, each on the even lane of the pin with the same
          // two bodies; the host launches this kernel only for such
          // schedules) run as one. Both lanes of a pair solve the pin alike
          // (same operands, so the same bits); the motor is then solved on
          // the pose the pin leaves (its own pose if the pin does not
          // commit), which is the pose the motor's round would read, and the
          // even lane writes the pair's two bodies once. Lambda stays in the
          // slot (both lanes of a pin pair keep it). The flags word is re-read
          // each iteration so that its per-round lane masks are formed where
          // they are used rather than held across the loop.
          static_assert(KX_REG_RTYPE == 0xAAA && RR == 12, "fused rounds: pins on even rounds, motors on odd");
          uint32_t fl = rflags;
          asm volatile("" : "+v"(fl));
          const bool odd = (sub & 1) != 0;
#pragma unroll
          for (int f = 0; f < RR / 2; f++) {
            KxRegSlot &c = slot[2 * f], &m = slot[2 * f + 1];
            // the fused rounds whose pins all have a second body (no grip):
            // a constant there, folded by the unrolled loop
            const bool livep = ((fl >> (2 * f)) & 1u) != 0u;
            const bool hasB = ((KX_REG_PINB >> f) & 1) != 0 || ((fl >> (16 + 2 * f)) & 1u) != 0u;
            const bool livem = ((fl >> (2 * f + 1)) & 1u) != 0u;
            const kx_f3v pA = *(const kx_f3v *)&pos[c.A], pB = *(const kx_f3v *)&pos[c.B];
            float s1, c1;
            kx_sincos(odd ? pB.z : pA.z, &s1, &c1);
            const float sa = kx_pair_even(s1), ca = kx_pair_even(c1), sb = kx_pair_odd(s1), cb = kx_pair_odd(c1);
            float xA = pA.x, yA = pA.y, tA = pA.z, xB = pB.x, yB = pB.y, tB = pB.z;
            float lx = c.lk.x, ly = c.lk.y;
            // (every lane runs both solves: the live flags only select)
            const bool okp = kx_pin_core_sc(xA, yA, tA, c.bA.x, c.bA.y, xB, yB, tB, c.bB.x, c.bB.y, hasB, c.cr.x,
                                            c.cr.y, c.dk.z, c.dk.w, c.dk.x, lx, ly, sa, ca, sb, cb) &&
                             livep;
            if (okp) c.lk = kx_f2v{lx, ly};
            xA = okp ? xA : pA.x; yA = okp ? yA : pA.y; tA = okp ? tA : pA.z;
            xB = okp ? xB : pB.x; yB = okp ? yB : pB.y; tB = okp ? tB : pB.z;
            float uA = tA, uB = tB, lam = m.lk.x;
            const bool okm =
                kx_angle_core_ab<true>(uA, uB, m.bA.y, m.bB.y, m.cr.x, m.dk.y, m.cr.y, m.cr.z, m.dk.x, m.dk.w, m.rm,
                                       lam) &&
                livem;
            if (okm) { tA = uA; tB = uB; m.lk.x = lam; }
            if (!odd && (okp || okm)) {
              // B first: for a world anchor B == A and pA must win
              *(kx_f3v *)&pos[c.B] = kx_f3v{xB, yB, tB};
              *(kx_f3v *)&pos[c.A] = kx_f3v{xA, yA, tA};
            }
            KX_WAVE_SYNC();
          }
        } else if constexpr (RR > 0) {
          // the fused rounds of the fixed schedule with the slots' records
          // read from LDS each round (launches of more than one wavefront per
          // SIMD, whose register file holds no slots)
          static_assert(KX_REG_RTYPE == 0xAAA && RR == 12, "fused rounds: pins on even rounds, motors on odd");
#pragma unroll
          for (int f = 0; f < RR / 2; f++)
            fused_pair(fdsc[2 * f], fdsc[2 * f + 1], ((KX_REG_PINB >> f) & 1) != 0, false);
        } else if (T.fused) {
          // a template the host marked fused (its rounds alternate pins and
          // motors, each motor on the lane of the pin before it with the
          // same two bodies): the fixed schedule's fused rounds with the
          // template's own round count; the slot descriptors of two fused
          // rounds come in one 16-byte LDS read
          const bool single = T.single != 0;
          for (int g = 0; 4 * g < T.R; g++) {
            const int4 d4 = act ? t_sched4[g * G + sub] : make_int4(-1, -1, -1, -1);
            fused_pair(d4.x, d4.y, false, single);
            if (4 * g + 2 < T.R) fused_pair(d4.z, d4.w, false, single);
          }
        } else {
        int4 d4n = act ? t_sched4[sub] : make_int4(-1, -1, -1, -1);
        // one round: its poses, then the next round's records into `nx`
        // (every load unconditional and ahead of the round-type branch, so
        // each join sees the same loads in flight), then the solve
        auto round = [&](const KxSlot &c, KxSlot &nx, int dnext, bool is_ang) {
          kx_f3v pA = *(const kx_f3v *)&pos[c.A], pB = *(const kx_f3v *)&pos[c.B];
          // the poses go to LDS first: the wait below is then for them alone
          __builtin_amdgcn_sched_barrier(0);
          nx = kx_slot_read(dnext, dyn, lam2, t_crec, t_brec);
          asm volatile("" : "+v"(pA), "+v"(pB));
          if (is_ang) {
            // crec {rest, min, max}, dyn {alpha, target, broken}, lambda,
            // angles, inverse inertias
            float tA = pA.z, tB = pB.z, lam = c.lk.x;
            const bool moved =
                kx_angle_core(tA, tB, c.bA.y, c.bB.y, c.hasB, c.cr.x, c.dk.y, c.cr.y, c.cr.z, c.dk.x, c.dk.w, lam);
            if (c.live && moved && __float_as_int(c.dk.z) == 0) {
              if (c.hasB) ((float *)&pos[c.B])[2] = tB;
              ((float *)&pos[c.A])[2] = tA;
              ((float *)&lam2[c.k])[0] = lam;
            }
          } else {
            // crec {lax, lay}, dyn {alpha, broken, lbx, lby}, lambda,
            // positions, {im, ii}; lanes 2j and 2j+1 hold the same pin:
            // each evaluates sin/cos of one body's angle and both lanes take
            // the pair's two results; the even lane commits
            const bool odd = (sub & 1) != 0;
            float s1, c1;
            kx_sincos(odd ? pB.z : pA.z, &s1, &c1);
            const float sa = kx_pair_even(s1), ca = kx_pair_even(c1), sb = kx_pair_odd(s1), cb = kx_pair_odd(c1);
            float xA = pA.x, yA = pA.y, tA = pA.z, xB = pB.x, yB = pB.y, tB = pB.z;
            float lx = c.lk.x, ly = c.lk.y;
            const bool moved = kx_pin_core_sc(xA, yA, tA, c.bA.x, c.bA.y, xB, yB, tB, c.bB.x, c.bB.y, c.hasB, c.cr.x,
                                              c.cr.y, c.dk.z, c.dk.w, c.dk.x, lx, ly, sa, ca, sb, cb) &&
                               !odd;
            if (c.live && moved && __float_as_int(c.dk.y) == 0) {
              // B first: for a world anchor B == A and pA must win
              *(kx_f3v *)&pos[c.B] = kx_f3v{xB, yB, tB};
              *(kx_f3v *)&pos[c.A] = kx_f3v{xA, yA, tA};
              lam2[c.k] = kx_f2v{lx, ly};
            }
          }
          KX_WAVE_SYNC();
        };
        // slot records alternate between two register sets: sa for the
        // even rounds of a group, sb for the odd ones
        KxSlot sa = kx_slot_read(d4n.x, dyn, lam2, t_crec, t_brec), sb;
        for (int r0 = 0; r0 < T.R; r0 += 4) {
          const uint32_t rbits = r0 < 32 ? rb0 : r0 < 64 ? rb1 : (uint32_t)__builtin_amdgcn_readfirstlane(t_rbits[r0 >> 5]);
          const int4 d4 = d4n;
          if (act && r0 + 4 < T.R) d4n = t_sched4[((r0 >> 2) + 1) * G + sub];
          round(sa, sb, d4.y, (rbits >> (r0 & 31)) & 1u);
          if (r0 + 1 >= T.R) break;
          round(sb, sa, d4.z, (rbits >> ((r0 + 1) & 31)) & 1u);
          if (r0 + 2 >= T.R) break;
          round(sa, sb, d4.w, (rbits >> ((r0 + 2) & 31)) & 1u);
          if (r0 + 3 >= T.R) break;
          round(sb, sa, d4n.x, (rbits >> ((r0 + 3) & 31)) & 1u);
        }
        }
        // collisions (xpbd.js:441-470: per body, its points in order, each
        // correction moving the body before the next point is tested).
        //
        // Lane `sub` owns bodies sub, sub + G, ...: a broad phase finds the
        // bodies that may touch a rect, and the owning lane folds each of
        // them; bodies are independent of one another here.
        {
          // A body is a candidate when its centre's cell is set in the near
          // bitmap: one LDS read per body. A clear cell lies farther than
          // reach + rmax (+ a margin for the binary32 cell expression) from
          // every rect's box, reach bounding every point's distance from its
          // body centre and rmax every point's radius, so no point of such a
          // body corrects; a centre outside the grid clamps to the edge cell,
          // which is no farther from any rect than the centre. A body with a
          // non-finite pose may be dropped too: a non-finite query point
          // never hits (the fold's distance stays infinite or NaN).
          // The bitmaps are read from LDS only; a map whose bitmaps do not
          // fit there gets no broad phase (every body is a candidate, which
          // is exact). A pointer that may point to LDS or to HBM would compile
          // to flat loads, which wait on both the LDS and the memory counters.
          const bool use_bm = P.bm_words != 0;
          const uint32_t *nearb = bits + (use_bm ? d.cell_words : 0);   // word 0: any word of the LDS allocation
          // (x + ox) - gx0 as x + (ox - gx0): ox and gx0 are integers, so the
          // difference is exact and the cell expression rounds once less
          const float oxg = ox - d.gx0, oyg = oy - d.gy0;
          // the near-bitmap word and bit of a body centre (clamped into the
          // grid; NaN converts to cell 0)
          auto cell = [&](float x, float y, int &wi, int &bi) {
            const int ix = min(max(__float2int_rd((x + oxg) * d.inv_ns), 0), d.nnx - 1);
            const int iy = min(max(__float2int_rd((y + oyg) * d.inv_ns), 0), d.nny - 1);
            const int c = iy * d.nnx + ix;
            wi = use_bm ? c >> 5 : 0;
            bi = c & 31;
          };
          // a lane's bodies sub, sub + G, ...: poses and bitmap words all
          // read before any test (no branch per body)
          uint32_t mine = 0u;
          {
            kx_f2v pbs[NBL];
#pragma unroll
            for (int k = 0; k < NBL; k++) {
              const int i = sub + k * G;
              pbs[k] = *(const kx_f2v *)&pos[i < nb ? i : 0];
            }
            uint32_t wd[NBL];
            int bi[NBL];
#pragma unroll
            for (int k = 0; k < NBL; k++) {
              int wi;
              cell(pbs[k].x, pbs[k].y, wi, bi[k]);
              wd[k] = nearb[wi];
            }
            const uint32_t all = use_bm ? 0u : 1u;   // no bitmaps: every body
#pragma unroll
            for (int k = 0; k < NBL; k++) mine |= (((wd[k] >> bi[k]) | all) & 1u) << k;
            mine &= bp_mask;
          }
          if constexpr (RR == 0) {   // (the register-slot kernel batches every body)
            for (int k = NBL, i = sub + NBL * G; i < nb; k++, i += G) {
              const float4 br = t_brec[i];
              if (br.x == 0.0f || (__float_as_int(br.z) >> 16) <= 0) continue;
              const float2 pb = *(const float2 *)&pos[i];
              int wi, bi;
              cell(pb.x, pb.y, wi, bi);
              if (!use_bm || ((nearb[wi] >> bi) & 1u) != 0u) mine |= 1u << k;
            }
          }
          // Second stage, exact per body: the near bitmap bounds every body
          // with the largest contact radius of the template and a whole cell
          // of slack; a body whose own contact radius (brec.w: offset length
          // + radius over its points, with a rounding margin) is within the
          // cull radius is kept only if some rect's box lies within that
          // radius of its centre (kx_centre_near_lds). A body dropped here
          // has no point that corrects at its pose after the rounds, so its
          // in-order fold would leave it unmoved.
          // The batched bodies go together: their contact radii and cell
          // offsets in one LDS round trip, the first two rects of each cell
          // in a second (a cell rarely holds more).
          if (P.map_words != 0) {
            for (uint32_t m = mine; m; m &= m - 1u) {
              const int k = __builtin_ctz(m), i = sub + k * G;
              const float D = t_brec[i].w;
              const float2 pb = *(const float2 *)&pos[i];
              if (D <= d.cull_r && !kx_centre_near_lds(d, m_off, m_crect, pb.x, pb.y, ox, oy, D)) mine &= ~(1u << k);
            }
          }
          // Two exact ways to fold the near bodies, chosen per launch:
          //  - parallel (launches of at most one wave per SIMD, whose time is
          //    the slowest wave's): the points of the near bodies go to a
          //    per-creature candidate list; phase A tests them in parallel at
          //    their body's pose after the rounds and flags those that
          //    correct; up to a body's first flagged point the sequential
          //    pass sees exactly that pose, so phase B (the owning lane)
          //    starts there and re-tests the rest in order. The order of the
          //    list does not affect any result.
          //  - owner (launches of many waves per SIMD, whose time is the sum):
          //    the sequential fold of each near body (xpbd.js:447-469) by the
          //    lane that owns it: its points in order, each queried at the
          //    body's current pose and correcting it when it penetrates.
          //    Points go two at a time: both are queried at the pose before
          //    the pair, and if the first corrects, the second is queried
          //    again at the corrected pose, so the fold is exactly the
          //    one-point-at-a-time fold. No candidate list, flags or atomics.
          if (__any(mine != 0u)) {
            if constexpr (PAR) {
              // clear the candidate count, then one reservation in the
              // creature's candidate list for all of the lane's points (the
              // flags are zero here: cleared at launch and after each use)
              if (act && sub == 0) evw[1] = 0u;
              KX_WAVE_SYNC();
              int ntot = 0;
              for (uint32_t m = mine; m; m &= m - 1u) ntot += __float_as_int(t_brec[sub + __builtin_ctz(m) * G].z) >> 16;
              if (ntot) {
                int at = (int)atomicAdd(&evw[1], (uint32_t)ntot);
                for (uint32_t m = mine; m; m &= m - 1u) {
                  const int i = sub + __builtin_ctz(m) * G;
                  const int bp = __float_as_int(t_brec[i].z), cnt = bp >> 16;
                  for (int t = 0; t < cnt; t++) cand[at + t] = (uint32_t)((bp & 0xffff) + t) | ((uint32_t)i << 16);
                  at += cnt;
                }
              }
              KX_WAVE_SYNC();
              const int ncand = act ? (int)evw[1] : 0;
              bool flagged = false;
              for (int j = sub; j < ncand; j += G) {
                const uint32_t e = cand[j];
                const int p = (int)(e & 0xffffu);
                const float4 pt = t_prec[p], pb = pos[e >> 16];
                float sa, ca, wx, wy;
                kx_sincos(pb.z, &sa, &ca);
                kx_world_point(pb.x, pb.y, ca, sa, pt.x, pt.y, wx, wy);
                KxHit hh;
                if (kx_query(d, P, bits, m_off, m_crect, wx, wy, ox, oy, hh) && pt.z - hh.d > KX_SLACK_POS) {
                  atomicOr(&flg[p >> 5], 1u << (p & 31));
                  flagged = true;
                }
              }
              // phase B only when some point of the wave corrects (rare on
              // the open maps this fold is chosen for)
              if (__any(flagged)) {
              KX_WAVE_SYNC();
              for (uint32_t m = mine; m; m &= m - 1u) {
                const int i = sub + __builtin_ctz(m) * G;
                float4 br = t_brec[i];
                const int bp = __float_as_int(br.z);
                const int o = bp & 0xffff, cnt = bp >> 16;
                int pq = cnt;                              // first flagged point of the body
                for (int wi = o >> 5; (wi << 5) < o + cnt; wi++) {
                  uint32_t wd = flg[wi];
                  const int lo = o - (wi << 5), hi = o + cnt - (wi << 5);
                  if (lo > 0) wd &= ~0u << lo;
                  if (hi < 32) wd &= (1u << hi) - 1u;
                  if (wd) { pq = (wi << 5) + __builtin_ctz(wd) - o; break; }
                }
                if (pq >= cnt) continue;
                float4 pb = pos[i];
                float x = pb.x, y = pb.y, sa, ca;
                kx_sincos(pb.z, &sa, &ca);
                for (; pq < cnt; pq++) {
                  float4 pt = t_prec[o + pq];
                  float wx, wy;
                  kx_world_point(x, y, ca, sa, pt.x, pt.y, wx, wy);
                  KxHit hh;
                  if (!kx_query(d, P, bits, m_off, m_crect, wx, wy, ox, oy, hh)) continue;
                  const float pen = pt.z - hh.d;
                  if (pen > KX_SLACK_POS) {
                    float nx, ny;
                    kx_hit_normal(hh, nx, ny);
                    x += nx * pen; y += ny * pen;
                  }
                }
                pb.x = x; pb.y = y;
                pos[i] = pb;
              }
              // every lane has read the flags: back to zero for the next pass
              KX_WAVE_SYNC();
              const int nfw = (T.np + 31) >> 5;
              for (int wi = sub; wi < (act ? nfw : 0); wi += G) flg[wi] = 0u;
              }
            } else {
              const bool lds_map = P.map_words != 0;
              for (uint32_t m = mine; m; m &= m - 1u) {
                const int i = sub + __builtin_ctz(m) * G;
                const int bp = __float_as_int(t_brec[i].z), o = bp & 0xffff, cnt = bp >> 16;
                float4 pb = pos[i];
                float x = pb.x, y = pb.y, sa, ca;
                kx_sincos(pb.z, &sa, &ca);
                bool moved = false;
                auto correct = [&](const KxHit &hh, float pen) {
                  float nx, ny;
                  kx_hit_normal(hh, nx, ny);
                  x += nx * pen; y += ny * pen;
                  moved = true;
                };
                for (int t = 0; t < cnt; t += 2) {
                  const bool two = t + 1 < cnt;
                  const float4 p0 = t_prec[o + t], p1 = t_prec[o + (two ? t + 1 : t)];
                  float wx0, wy0, wx1, wy1;
                  kx_world_point(x, y, ca, sa, p0.x, p0.y, wx0, wy0);
                  kx_world_point(x, y, ca, sa, p1.x, p1.y, wx1, wy1);
                  bool q0, q1;
                  KxHit h0, h1;
                  if (lds_map) kx_sdf2_cells_lds(d, m_off, m_crect, wx0, wy0, wx1, wy1, ox, oy, q0, h0, q1, h1);
                  else kx_sdf2_cells_glb(d, P.bm_words ? bits : nullptr, wx0, wy0, wx1, wy1, ox, oy, q0, h0, q1, h1);
                  if (q0 && p0.z - h0.d > KX_SLACK_POS) {
                    correct(h0, p0.z - h0.d);
                    if (two) {
                      kx_world_point(x, y, ca, sa, p1.x, p1.y, wx1, wy1);
                      KxHit h;
                      if (kx_query(d, P, bits, m_off, m_crect, wx1, wy1, ox, oy, h) && p1.z - h.d > KX_SLACK_POS)
                        correct(h, p1.z - h.d);
                    }
                  } else if (two && q1 && p1.z - h1.d > KX_SLACK_POS) {
                    correct(h1, p1.z - h1.d);
                  }
                }
                if (moved) {
                  pb.x = x; pb.y = y;
                  pos[i] = pb;
                }
              }
            }
          }
        }
        KX_WAVE_SYNC();
      }
#pragma unroll
      for (int k = 0; k < BR; k++) {
        if (k >= nbl) break;
        const int i = sub + k * G;
        if (i >= nb) continue;
        float4 pp = pos[i];
        kx_finalize(d, rx[k], ry[k], ra[k], rvx[k], rvy[k], rva[k], pp.x, pp.y, pp.z);
        pos[i] = pp;
      }
      KX_WAVE_SYNC();
    }
    if constexpr (RR > 0 && REG) {
      // the last substep's lambdas back to the creature region (grip break,
      // store); a pin's even lane writes it
#pragma unroll
      for (int r = 0; r < RR; r++) {
        const int dsc = slot_dsc(r);
        const bool is_ang = ((KX_REG_RTYPE >> r) & 1) != 0;
        if (dsc >= 0 && (is_ang || (sub & 1) == 0)) lam2[dsc & 255] = slot[r].lk;
      }
      KX_WAVE_SYNC();
    }
    // grip break (xpbd.js:496-504 / solver.wgsl:433-442)
    for (int j = sub; j < (act ? T.ngrip : 0); j += G) {
      const int k = t_grip_k[j];
      if (__float_as_int(((const float *)&dyn[k])[1])) continue;
      const kx_f2v lk = lam2[k];
      float f = sqrtf(lk.x * lk.x + lk.y * lk.y) * d.invH;
      if (f > gbf[j]) {
        lam2[k] = kx_f2v{0.0f, 0.0f};
        ((float *)&dyn[k])[1] = __int_as_float(1);
        int arm = t_grip_arm[j];
        if (arm >= 0) my_ev |= 4u << (3 * arm);
      }
    }
    if (d.events) {
      if (my_ev) atomicOr(evw, my_ev);
      KX_WAVE_SYNC();
      if (act && sub == 0) d.events[(size_t)(ev_row0 + step) * d.nCr + cid] = evw[0];
    }
#pragma unroll
    for (int k = 0; k < BR; k++) {
      if (k >= nbl) break;
      const int i = sub + k * G;
      rx[k] += ox; ry[k] += oy;
      if (i < nb) {
        float4 pp = pos[i];
        pp.x += ox; pp.y += oy;
        pos[i] = pp;
      }
    }
    kx_shift_anchors(dyn, t_crec, nc, sub, G, wa_mask, ox, oy);
    publish_tips();
    KX_WAVE_SYNC();
  }

  // ── store ──
#pragma unroll
  for (int k = 0; k < BR; k++) {
    if (k >= nbl) break;
    const int i = sub + k * G;
    if (i >= nb) continue;
    const int g = b0 + i;
    const float4 pp = pos[i];
    d.px[g] = rx[k]; d.py[g] = ry[k]; d.ang[g] = ra[k];
    d.vx[g] = rvx[k]; d.vy[g] = rvy[k]; d.av[g] = rva[k];
    d.ppx[g] = pp.x; d.ppy[g] = pp.y; d.angP[g] = pp.z;
  }
  for (int k = sub; k < nc; k += G) {
    const int g = c0 + k;
    const int v = __float_as_int(t_crec[k].w);
    const float4 dk = dyn[k];
    const kx_f2v lk = lam2[k];
    if ((v >> 16) == KX_ANGLE) {
      d.broken[g] = __float_as_int(dk.z);
      d.target[g] = dk.y;
      if (has_law) d.comp[g] = (float)comp_last;
      d.lam[g] = lk.x; d.lamx[g] = 0.0f; d.lamy[g] = 0.0f;
    } else {
      d.broken[g] = __float_as_int(dk.y);
      if (((v >> 8) & 255) == 255) { d.lbx[g] = dk.z; d.lby[g] = dk.w; }
      d.lamx[g] = lk.x; d.lamy[g] = lk.y; d.lam[g] = 0.0f;
    }
  }
  for (int j = sub; j < (act ? T.ngrip : 0); j += G) d.bf[c0 + t_grip_k[j]] = gbf[j];
  if (act && sub < T.narms) d.cd[cid * KX_MAX_ARMS + sub] = my_cd;
}

// Every workgroup with the same lane count G; PAR: the collision fold
// (P.coll_par), one kernel per fold so that each compiles alone.
template <int G, int WPB, bool PAR>
__global__ __launch_bounds__(KX_WAVE *WPB) __attribute__((amdgpu_waves_per_eu(2))) void kx_k_persist(KxDev d, KxLaw L, int has_law, int step0,
                                                             const int *step_base, int nsteps, int ev_row0,
                                                             int with_forces, KxPersistCfg P) {
  kx_persist_body<G, WPB, PAR>(d, L, has_law, step0, step_base, nsteps, ev_row0, with_forces, P, blockIdx.x);
}

// The same with RR register-held round slots per lane: at most one
// wavefront per SIMD (the slots take the register file), for launches that
// place no more than that.
template <int G, int WPB, bool PAR, int RR>
__global__ __launch_bounds__(KX_WAVE *WPB) __attribute__((amdgpu_waves_per_eu(1, 1))) void kx_k_persist_reg(
    KxDev d, KxLaw L, int has_law, int step0, const int *step_base, int nsteps, int ev_row0, int with_forces,
    KxPersistCfg P) {
  kx_persist_body<G, WPB, PAR, RR>(d, L, has_law, step0, step_base, nsteps, ev_row0, with_forces, P, blockIdx.x);
}

// The fixed schedule with the slots read from LDS: launches of the same
// templates with more than one wavefront per SIMD.
template <int G, int WPB, bool PAR, int RR>
__global__ __launch_bounds__(KX_WAVE *WPB) __attribute__((amdgpu_waves_per_eu(2))) void kx_k_persist_fix(
    KxDev d, KxLaw L, int has_law, int step0, const int *step_base, int nsteps, int ev_row0, int with_forces,
    KxPersistCfg P) {
  kx_persist_body<G, WPB, PAR, RR, false>(d, L, has_law, step0, step_base, nsteps, ev_row0, with_forces, P,
                                          blockIdx.x);
}

// Lane count per workgroup (per morphology bucket): a wave-uniform switch to
// the body instantiated for that count. Every persistent kernel is held to
// 256 VGPRs (two waves per SIMD): no launch here has LDS for more.
template <int WPB, bool PAR>
__global__ __launch_bounds__(KX_WAVE *WPB) __attribute__((amdgpu_waves_per_eu(2))) void kx_k_persist_mix(KxDev d, KxLaw L, int has_law, int step0,
                                                                 const int *step_base, int nsteps, int ev_row0,
                                                                 int with_forces, KxPersistCfg P) {
  const int bid = blockIdx.x;
  switch (P.wg_lanes[bid]) {
    case 8: kx_persist_body<8, WPB, PAR>(d, L, has_law, step0, step_base, nsteps, ev_row0, with_forces, P, bid); break;
    case 16: kx_persist_body<16, WPB, PAR>(d, L, has_law, step0, step_base, nsteps, ev_row0, with_forces, P, bid); break;
    case 32: kx_persist_body<32, WPB, PAR>(d, L, has_law, step0, step_base, nsteps, ev_row0, with_forces, P, bid); break;
    default: kx_persist_body<64, WPB, PAR>(d, L, has_law, step0, step_base, nsteps, ev_row0, with_forces, P, bid); break;
  }
}

// A pointer read from memory is generic (flat) to the compiler; taking it as
// a global pointer lets every access through it compile to a global load or
// store, which waits on the memory counter only (a flat access also holds
// back the LDS counter). The cast changes no bits; the empty asm keeps the
// compiler from folding the cast pair away.
template <class T>
__device__ __forceinline__ T *kx_glob(T *p) {
  __attribute__((address_space(1))) T *g = (__attribute__((address_space(1))) T *)p;
  asm volatile("" : "+s"(g));
  return (T *)g;
}

// Several handles in one grid (the serving loop's continuous batch): each
// workgroup runs its entry's body with that entry's state, tables, law and
// step count. The entry's HBM pointers are taken as global pointers.
template <int WPB, bool PAR>
__global__ __launch_bounds__(KX_WAVE *WPB) __attribute__((amdgpu_waves_per_eu(2))) void kx_k_persist_batch(const KxBatchEnt *__restrict__ E,
                                                                   const int2 *__restrict__ wgmap) {
  const int2 m = wgmap[blockIdx.x];
  const KxBatchEnt &e = E[m.x];
  const int bid = m.y;
  KxDev d = e.d;
  d.px = kx_glob(d.px); d.py = kx_glob(d.py); d.ang = kx_glob(d.ang);
  d.vx = kx_glob(d.vx); d.vy = kx_glob(d.vy); d.av = kx_glob(d.av);
  d.ppx = kx_glob(d.ppx); d.ppy = kx_glob(d.ppy); d.angP = kx_glob(d.angP);
  d.fx = kx_glob(d.fx); d.fy = kx_glob(d.fy); d.tq = kx_glob(d.tq);
  d.comp = kx_glob(d.comp); d.target = kx_glob(d.target); d.lbx = kx_glob(d.lbx); d.lby = kx_glob(d.lby);
  d.bf = kx_glob(d.bf); d.broken = kx_glob(d.broken);
  d.lam = kx_glob(d.lam); d.lamx = kx_glob(d.lamx); d.lamy = kx_glob(d.lamy);
  d.rects = kx_glob(d.rects); d.cell_off = kx_glob(d.cell_off); d.cell_rects = kx_glob(d.cell_rects);
  d.cell_bits = kx_glob(d.cell_bits); d.cr = kx_glob(d.cr); d.cd = kx_glob(d.cd); d.tape = kx_glob(d.tape);
  d.events = nullptr;
  KxPersistCfg P = e.P;
  P.wave_cr = kx_glob(P.wave_cr); P.wg_tmpl = kx_glob(P.wg_tmpl); P.wg_lanes = kx_glob(P.wg_lanes);
  P.wg_cr_off = kx_glob(P.wg_cr_off); P.tmpl = kx_glob(P.tmpl); P.pool = kx_glob(P.pool);
  const KxLaw L = e.L;
  switch (P.wg_lanes[bid]) {
    case 8: kx_persist_body<8, WPB, PAR>(d, L, e.has_law, e.step0, nullptr, e.nsteps, 0, 0, P, bid); break;
    case 16: kx_persist_body<16, WPB, PAR>(d, L, e.has_law, e.step0, nullptr, e.nsteps, 0, 0, P, bid); break;
    case 32: kx_persist_body<32, WPB, PAR>(d, L, e.has_law, e.step0, nullptr, e.nsteps, 0, 0, P, bid); break;
    default: kx_persist_body<64, WPB, PAR>(d, L, e.has_law, e.step0, nullptr, e.nsteps, 0, 0, P, bid); break;
  }
}

// Raises a kernel's dynamic LDS limit to 160 KiB once per device; safe to
// call from several threads at once (the attribute call is idempotent and
// the per-device flags are atomic).
static void allow_full_lds(const void *fn, std::atomic<uint64_t> &done) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
  const uint64_t bit = 1ull << dev;
  if (done.load(std::memory_order_acquire) & bit) return;
  (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
  done.fetch_or(bit, std::memory_order_acq_rel);
}

template <int WPB, bool PAR>
static int launch_batch(const KxBatchEnt *E, const int2 *wgmap, int nwg, size_t lds, hipStream_t st) {
  static std::atomic<uint64_t> lds_set{0};
  allow_full_lds((const void *)kx_k_persist_batch<WPB, PAR>, lds_set);
  hipLaunchKernelGGL((kx_k_persist_batch<WPB, PAR>), dim3(nwg), dim3(KX_WAVE * WPB), lds, st, E, wgmap);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

int kx_persist_launch_batch(int wpb, int par, const KxBatchEnt *ents_dev, const int2 *wgmap_dev, int nwg, size_t lds,
                            hipStream_t st) {
  if (nwg <= 0) return 0;
  if (lds > 160 * 1024) return -22;
  switch (wpb * 2 + (par ? 1 : 0)) {
    case 2: return launch_batch<1, false>(ents_dev, wgmap_dev, nwg, lds, st);
    case 3: return launch_batch<1, true>(ents_dev, wgmap_dev, nwg, lds, st);
    case 4: return launch_batch<2, false>(ents_dev, wgmap_dev, nwg, lds, st);
    case 5: return launch_batch<2, true>(ents_dev, wgmap_dev, nwg, lds, st);
    default: return -22;
  }
}

__global__ void kx_k_bump(int *counter, int delta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += delta;
}

int kx_bump_counter(int *counter, int delta, hipStream_t st) {
  hipLaunchKernelGGL(kx_k_bump, dim3(1), dim3(1), 0, st, counter, delta);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <int G, int WPB, bool PAR>
static int launch_gw(const KxDev &d, const KxLaw *law, int step0, const int *step_base, int nsteps, int ev_row0,
                     int with_forces, const KxPersistCfg &P, hipStream_t st) {
  size_t lds = ((size_t)P.tmpl_words + P.bm_words + P.map_words + (size_t)P.cr_words) * 4;
  if (lds > 160 * 1024) return -22;
  static std::atomic<uint64_t> lds_set{0};
  allow_full_lds((const void *)kx_k_persist<G, WPB, PAR>, lds_set);
  KxLaw L = law ? *law : KxLaw{};
  hipLaunchKernelGGL((kx_k_persist<G, WPB, PAR>), dim3(P.nwg), dim3(KX_WAVE * WPB), lds, st, d, L, law ? 1 : 0,
                     step0, step_base, nsteps, ev_row0, with_forces, P);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <int G, int WPB, bool PAR, int RR>
static int launch_reg(const KxDev &d, const KxLaw *law, int step0, const int *step_base, int nsteps, int ev_row0,
                      int with_forces, const KxPersistCfg &P, hipStream_t st) {
  size_t lds = ((size_t)P.tmpl_words + P.bm_words + P.map_words + (size_t)P.cr_words) * 4;
  if (lds > 160 * 1024) return -22;
  static std::atomic<uint64_t> lds_set{0};
  allow_full_lds((const void *)kx_k_persist_reg<G, WPB, PAR, RR>, lds_set);
  KxLaw L = law ? *law : KxLaw{};
  hipLaunchKernelGGL((kx_k_persist_reg<G, WPB, PAR, RR>), dim3(P.nwg), dim3(KX_WAVE * WPB), lds, st, d, L,
                     law ? 1 : 0, step0, step_base, nsteps, ev_row0, with_forces, P);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <int G, int WPB, bool PAR, int RR>
static int launch_fix(const KxDev &d, const KxLaw *law, int step0, const int *step_base, int nsteps, int ev_row0,
                      int with_forces, const KxPersistCfg &P, hipStream_t st) {
  size_t lds = ((size_t)P.tmpl_words + P.bm_words + P.map_words + (size_t)P.cr_words) * 4;
  if (lds > 160 * 1024) return -22;
  static std::atomic<uint64_t> lds_set{0};
  allow_full_lds((const void *)kx_k_persist_fix<G, WPB, PAR, RR>, lds_set);
  KxLaw L = law ? *law : KxLaw{};
  hipLaunchKernelGGL((kx_k_persist_fix<G, WPB, PAR, RR>), dim3(P.nwg), dim3(KX_WAVE * WPB), lds, st, d, L,
                     law ? 1 : 0, step0, step_base, nsteps, ev_row0, with_forces, P);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <int G, int WPB, bool PAR, int RR>
static int occ_reg(size_t lds) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)kx_k_persist_reg<G, WPB, PAR, RR>, KX_WAVE * WPB,
                                                   lds) != hipSuccess)
    return 0;
  return n;
}

// Resident workgroups per CU of the register-slot kernel (16 lanes, 12
// rounds) for wpb and fold `par`, 0 when not instantiated.
int kx_persist_occupancy_reg(int wpb, size_t lds, bool par) {
  if (par) return wpb == 1 ? occ_reg<16, 1, true, KX_REG_ROUNDS>(lds) : wpb == 2 ? occ_reg<16, 2, true, KX_REG_ROUNDS>(lds)
                : wpb == 4 ? occ_reg<16, 4, true, KX_REG_ROUNDS>(lds) : 0;
  return wpb == 1 ? occ_reg<16, 1, false, KX_REG_ROUNDS>(lds) : wpb == 2 ? occ_reg<16, 2, false, KX_REG_ROUNDS>(lds)
         : wpb == 4 ? occ_reg<16, 4, false, KX_REG_ROUNDS>(lds) : 0;
}

template <int WPB, bool PAR>
static int launch_mix(const KxDev &d, const KxLaw *law, int step0, const int *step_base, int nsteps, int ev_row0,
                      int with_forces, const KxPersistCfg &P, hipStream_t st) {
  size_t lds = ((size_t)P.tmpl_words + P.bm_words + P.map_words + (size_t)P.cr_words) * 4;
  if (lds > 160 * 1024) return -22;
  static std::atomic<uint64_t> lds_set{0};
  allow_full_lds((const void *)kx_k_persist_mix<WPB, PAR>, lds_set);
  KxLaw L = law ? *law : KxLaw{};
  hipLaunchKernelGGL((kx_k_persist_mix<WPB, PAR>), dim3(P.nwg), dim3(KX_WAVE * WPB), lds, st, d, L, law ? 1 : 0,
                     step0, step_base, nsteps, ev_row0, with_forces, P);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <int WPB, bool PAR>
static int occ_mix(size_t lds) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)kx_k_persist_mix<WPB, PAR>, KX_WAVE * WPB,
                                                   lds) != hipSuccess)
    return 0;
  return n;
}

template <int G, int WPB, bool PAR>
static int occ_gw(size_t lds) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, (const void *)kx_k_persist<G, WPB, PAR>, KX_WAVE * WPB,
                                                   lds) != hipSuccess)
    return 0;
  return n;
}

template <bool PAR>
static int occupancy(int G, int wpb, size_t lds) {
  if (G == 0)
    return wpb == 1 ? occ_mix<1, PAR>(lds) : wpb == 2 ? occ_mix<2, PAR>(lds) : wpb == 4 ? occ_mix<4, PAR>(lds) : 0;
#define KX_OCC(g)                                   \
  case g:                                           \
    return wpb == 1 ? occ_gw<g, 1, PAR>(lds)        \
         : wpb == 2 ? occ_gw<g, 2, PAR>(lds)        \
         : wpb == 4 ? occ_gw<g, 4, PAR>(lds) : 0;
  switch (G) { KX_OCC(8) KX_OCC(16) KX_OCC(32) KX_OCC(64) default: return 0; }
#undef KX_OCC
}

// Resident workgroups per CU for shape (G, wpb) and collision fold `par` with
// `lds` bytes of dynamic LDS each (0 when the shape is not instantiated or
// the query fails).
int kx_persist_occupancy(int G, int wpb, size_t lds, bool par) {
  return par ? occupancy<true>(G, wpb, lds) : occupancy<false>(G, wpb, lds);
}

template <int G, bool PAR>
static int launch_g(int wpb, const KxDev &d, const KxLaw *law, int step0, const int *step_base, int nsteps,
                    int ev_row0, int with_forces, const KxPersistCfg &P, hipStream_t st) {
  if constexpr (G == 16) {
    if (P.reg_rounds == KX_REG_ROUNDS) {
      switch (wpb) {
        case 1: return launch_reg<16, 1, PAR, KX_REG_ROUNDS>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
        case 2: return launch_reg<16, 2, PAR, KX_REG_ROUNDS>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
        case 4: return launch_reg<16, 4, PAR, KX_REG_ROUNDS>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
        default: return -22;
      }
    }
    if (P.fixed_rounds == KX_REG_ROUNDS) {
      switch (wpb) {
        case 1: return launch_fix<16, 1, PAR, KX_REG_ROUNDS>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
        case 2: return launch_fix<16, 2, PAR, KX_REG_ROUNDS>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
        case 4: return launch_fix<16, 4, PAR, KX_REG_ROUNDS>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
        default: return -22;
      }
    }
  }
  switch (wpb) {
    case 1: return launch_gw<G, 1, PAR>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 2: return launch_gw<G, 2, PAR>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 4: return launch_gw<G, 4, PAR>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    default: return -22;
  }
}

template <bool PAR>
static int launch_par(int G, int wpb, const KxDev &d, const KxLaw *law, int step0, const int *step_base,
                      int nsteps, int ev_row0, int with_forces, const KxPersistCfg &P, hipStream_t st) {
  if (P.mixed) {
    switch (wpb) {
      case 1: return launch_mix<1, PAR>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
      case 2: return launch_mix<2, PAR>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
      case 4: return launch_mix<4, PAR>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
      default: return -22;
    }
  }
  switch (G) {
    case 8: return launch_g<8, PAR>(wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 16: return launch_g<16, PAR>(wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 32: return launch_g<32, PAR>(wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 64: return launch_g<64, PAR>(wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    default: return -22;
  }
}

int kx_persist_launch(int G, int wpb, const KxDev &d, const KxLaw *law, int step0, const int *step_base,
                      int nsteps, int ev_row0, int with_forces, const KxPersistCfg &P, hipStream_t st) {
  return P.coll_par ? launch_par<true>(G, wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st)
                    : launch_par<false>(G, wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
}
