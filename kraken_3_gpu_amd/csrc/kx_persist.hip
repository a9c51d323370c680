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
// Positions live in creature-local coordinates relative to an integer
// origin (re-centred on the creature's first body between steps); since a
// global binary32 coordinate is a multiple of its own ulp, the shift to and
// from the origin is exact and the solver works with finer position ulps.
#include "../../include/kxpbd.h"
#include "kx_device.h"
#include "kx_sched.h"

#define KX_WAVE_SYNC()                                     \
  do {                                                     \
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
    __builtin_amdgcn_wave_barrier();                       \
  } while (0)

// per-creature LDS layout (word offsets, 16-byte records):
//   pos[i] = {ppx, ppy, angP, -}   prv[i] = {px, py, ang, -}   vel[i] = {vx, vy, av, -}
//   dyn[k] = {alpha, target, lbx, lby}   lam[k] = {lambda_x|lambda, lambda_y, breakForce, broken}
struct KxCrLds {
  int pos, prv, vel, dyn, lam, ev, words;
  __host__ __device__ KxCrLds(int NB, int NC) {
    int o = 0;
    pos = o; o += 4 * NB; prv = o; o += 4 * NB; vel = o; o += 4 * NB;
    dyn = o; o += 4 * NC; lam = o; o += 4 * NC;
    ev = o; o += 4;
    words = o;
  }
};

int kx_persist_cr_words(int NB, int NC) { return KxCrLds(NB, NC).words; }

template <int G, int WPB>
__global__ __launch_bounds__(KX_WAVE *WPB) void kx_k_persist(KxDev d, KxLaw L, int has_law, int step0,
                                                             const int *step_base, int nsteps, int ev_row0,
                                                             int with_forces, KxPersistCfg P) {
  constexpr int CPW = KX_WAVE / G;
  extern __shared__ __attribute__((aligned(16))) uint32_t smem[];
  const int lane = threadIdx.x & (KX_WAVE - 1);
  const int q = lane / G, sub = lane % G;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / KX_WAVE);
  const int gw = blockIdx.x * WPB + w;

  // ── template + cell bitmap -> LDS (the only workgroup barrier) ──
  const int tid = P.wave_tmpl[blockIdx.x * WPB];
  const KxTmpl T = P.tmpl[tid];
  for (int i = threadIdx.x; i < T.words; i += blockDim.x) smem[i] = P.pool[T.base + i];
  uint32_t *bits = smem + P.tmpl_words;
  for (int i = threadIdx.x; i < P.bm_words; i += blockDim.x) bits[i] = d.cell_bits[i];
  __syncthreads();
  if (gw >= P.nwaves) return;

  const float4 *t_crec = (const float4 *)(smem + T.o_crec);   // 2 per constraint
  const float4 *t_brec = (const float4 *)(smem + T.o_brec);
  const float4 *t_prec = (const float4 *)(smem + T.o_prec);
  const int *t_sched = (const int *)(smem + T.o_sched), *t_rtype = (const int *)(smem + T.o_rtype);
  const int *t_coll = (const int *)(smem + T.o_coll), *t_ccnt = (const int *)(smem + T.o_ccnt);
  const int *t_arm_tip = (const int *)(smem + T.o_arm_tip), *t_arm_grip = (const int *)(smem + T.o_arm_grip);
  const int *t_mot_k = (const int *)(smem + T.o_mot_k), *t_mot_slot = (const int *)(smem + T.o_mot_slot);
  const int *t_grip_k = (const int *)(smem + T.o_grip_k), *t_grip_arm = (const int *)(smem + T.o_grip_arm);

  const int cid = P.wave_cr[gw * CPW + q];
  const bool act = cid >= 0;
  const KxCrLds Lo(P.NB, P.NC);
  uint32_t *S = smem + P.tmpl_words + P.bm_words + (size_t)(w * CPW + q) * P.cr_stride;
  float4 *pos = (float4 *)(S + Lo.pos), *prv = (float4 *)(S + Lo.prv), *vel = (float4 *)(S + Lo.vel);
  float4 *dyn = (float4 *)(S + Lo.dyn), *lamr = (float4 *)(S + Lo.lam);
  uint32_t *evw = S + Lo.ev;
  const int nb = act ? T.nb : 0, nc = act ? T.nc : 0;
  const int b0 = act ? d.cr[cid].body_off : 0, c0 = act ? d.cr[cid].cons_off : 0;

  float ox = 0.0f, oy = 0.0f;
  if (act) { ox = rintf(d.px[b0]); oy = rintf(d.py[b0]); }

  // ── load the creature's dynamic state ──
  for (int i = sub; i < nb; i += G) {
    int g = b0 + i;
    prv[i] = make_float4(d.px[g] - ox, d.py[g] - oy, d.ang[g], 0.0f);
    vel[i] = make_float4(d.vx[g], d.vy[g], d.av[g], 0.0f);
  }
  for (int k = sub; k < nc; k += G) {
    int g = c0 + k;
    bool world = ((__float_as_int(t_crec[2 * k].x) >> 8) & 255) == 255;
    dyn[k] = make_float4(d.comp[g] * d.invH2, d.target[g], world ? d.lbx[g] - ox : d.lbx[g],
                         world ? d.lby[g] - oy : d.lby[g]);
    lamr[k] = make_float4(0.0f, 0.0f, d.bf[g], __int_as_float(d.broken[g]));
  }
  double my_cd = 0.0;
  if (act && sub < T.narms) my_cd = d.cd[cid * KX_MAX_ARMS + sub];
  const double hS = d.hd * d.S;
  KX_WAVE_SYNC();

  for (int step = 0; step < nsteps; step++) {
    const int key = (step_base ? step_base[0] : step0) + step;
    if (act && sub == 0) evw[0] = 0u;
    uint32_t my_ev = 0u;
    if (has_law) {
      // motor targets + compliance (core.js:190-198)
      const float a_law = (float)L.comp * d.invH2;
      for (int j = sub; j < (act ? T.nmot : 0); j += G) {
        int k = t_mot_k[j];
        float4 dk = dyn[k];
        dk.x = a_law;
        dk.y = kx_law_target(L, (uint32_t)cid, (uint32_t)key, t_mot_slot[j]);
        dyn[k] = dk;
      }
      // adhesion reflex (core.js:199-218), one lane per arm
      if (act && sub < T.narms) {
        double adh = kx_ctl_u(L.seed, (uint32_t)cid, (uint32_t)key, KX_SLOT_ADH + sub);
        int g = t_arm_grip[sub], t = t_arm_tip[sub];
        float4 dk = dyn[g], lk = lamr[g], tp = prv[t], tv = vel[t], cr0 = t_crec[2 * g];
        int bk = __float_as_int(lk.w);
        float f_bf = lk.z, gx = dk.z + ox, gy = dk.w + oy, lx = lk.x, ly = lk.y;
        uint32_t e = kx_reflex(L, hS, adh, my_cd, bk, f_bf, gx, gy, lx, ly, tv.x, tv.y, (double)ox + (double)tp.x,
                               (double)oy + (double)tp.y, tp.z, cr0.y, cr0.z);
        dk.z = gx - ox; dk.w = gy - oy;
        dyn[g] = dk;
        lamr[g] = make_float4(lx, ly, f_bf, __int_as_float(bk));
        my_ev |= e << (3 * sub);
      }
      KX_WAVE_SYNC();
    }
    for (int s = 0; s < d.S; s++) {
      // resetLambda + integrate (xpbd.js:262-287)
      for (int k = sub; k < nc; k += G) { float4 lk = lamr[k]; lk.x = 0.0f; lk.y = 0.0f; lamr[k] = lk; }
      const bool forces = with_forces && step == 0 && s == 0;
      for (int i = sub; i < nb; i += G) {
        float fx = 0.0f, fy = 0.0f, tq = 0.0f;
        if (forces) {
          int g = b0 + i;
          fx = d.fx[g]; fy = d.fy[g]; tq = d.tq[g];
          d.fx[g] = 0.0f; d.fy[g] = 0.0f; d.tq[g] = 0.0f;
        }
        float4 p = prv[i], v = vel[i], br = t_brec[i];
        float q_x, q_y, q_a;
        kx_integrate(d, v.x, v.y, v.z, p.x, p.y, p.z, br.x, br.y, fx, fy, tq, q_x, q_y, q_a);
        vel[i] = v;
        pos[i] = make_float4(q_x, q_y, q_a, 0.0f);
      }
      KX_WAVE_SYNC();
      for (int it = 0; it < d.I; it++) {
        // constraint rounds (type-homogeneous, wave-uniform branch)
        for (int r = 0; r < T.R; r++) {
          const int k = act ? t_sched[r * G + sub] : -1;
          if (k >= 0) {
            float4 lk = lamr[k];
            if (__float_as_int(lk.w) == 0) {
              float4 c0r = t_crec[2 * k];
              int v = __float_as_int(c0r.x);
              int A = v & 255, B = (v >> 8) & 255;
              bool hasB = B != 255;
              if (!hasB) B = A;
              float4 pA = pos[A], pB = pos[B];
              float4 bA = t_brec[A], bB = t_brec[B];
              float4 dk = dyn[k];
              if (t_rtype[r]) {
                float4 c1r = t_crec[2 * k + 1];
                if (kx_angle_core(pA.z, pB.z, bA.y, bB.y, hasB, c0r.w, dk.y, c1r.x, c1r.y, dk.x, lk.x)) {
                  pos[A] = pA;
                  if (hasB) pos[B] = pB;
                  lamr[k] = lk;
                }
              } else {
                if (kx_pin_core(pA.x, pA.y, pA.z, bA.x, bA.y, pB.x, pB.y, pB.z, bB.x, bB.y, hasB, c0r.y, c0r.z,
                                dk.z, dk.w, dk.x, lk.x, lk.y)) {
                  pos[A] = pA;
                  if (hasB) pos[B] = pB;
                  lamr[k] = lk;
                }
              }
            }
          }
          KX_WAVE_SYNC();
        }
        // collisions: this lane's bodies, each body's points in order
        const int nbl = act ? t_ccnt[sub] : 0;
        for (int j = 0; j < nbl; j++) {
          int i = t_coll[sub * T.maxbl + j];
          float4 br = t_brec[i];
          if (br.x == 0.0f) continue;
          int bp = __float_as_int(br.z);
          int o = bp & 0xffff, cnt = bp >> 16;
          float4 p = pos[i];
          float x = p.x, y = p.y;
          float sa, ca;
          kx_sincos(p.z, &sa, &ca);
          for (int pq = 0; pq < cnt; pq++) {
            float4 pt = t_prec[o + pq];
            float wx = x + (pt.x * ca - pt.y * sa);
            float wy = y + (pt.x * sa + pt.y * ca);
            KxHit hh;
            bool hit = P.bm_words ? kx_sdf_grid_bm(d, bits, wx, wy, ox, oy, hh) : kx_sdf_grid(d, wx, wy, ox, oy, hh);
            if (!hit) continue;
            float pen = pt.z - hh.d;
            if (pen > KX_SLACK_POS) { x += hh.nx * pen; y += hh.ny * pen; }
          }
          p.x = x; p.y = y;
          pos[i] = p;
        }
        KX_WAVE_SYNC();
      }
      for (int i = sub; i < nb; i += G) {
        float4 p = prv[i], pp = pos[i];
        float vx, vy, av;
        kx_finalize(d, p.x, p.y, p.z, vx, vy, av, pp.x, pp.y, pp.z);
        prv[i] = p;
        pos[i] = pp;
        vel[i] = make_float4(vx, vy, av, 0.0f);
      }
      KX_WAVE_SYNC();
    }
    // grip break (xpbd.js:496-504 / solver.wgsl:433-442)
    for (int j = sub; j < (act ? T.ngrip : 0); j += G) {
      int k = t_grip_k[j];
      float4 lk = lamr[k];
      if (__float_as_int(lk.w)) continue;
      float f = sqrtf(lk.x * lk.x + lk.y * lk.y) * d.invH;
      if (f > lk.z) {
        lamr[k] = make_float4(0.0f, 0.0f, lk.z, __int_as_float(1));
        int arm = t_grip_arm[j];
        if (arm >= 0) my_ev |= 4u << (3 * arm);
      }
    }
    if (d.events) {
      if (my_ev) atomicOr(evw, my_ev);
      KX_WAVE_SYNC();
      if (act && sub == 0) d.events[(size_t)(ev_row0 + step) * d.nCr + cid] = evw[0];
    }
    // re-centre the origin on the first body when it drifted
    float4 p0 = act ? prv[0] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float sx = rintf(p0.x), sy = rintf(p0.y);
    KX_WAVE_SYNC();
    if (fabsf(sx) > 32.0f || fabsf(sy) > 32.0f) {
      for (int i = sub; i < nb; i += G) {
        float4 p = prv[i], pp = pos[i];
        p.x -= sx; p.y -= sy; pp.x -= sx; pp.y -= sy;
        prv[i] = p; pos[i] = pp;
      }
      for (int k = sub; k < nc; k += G)
        if (((__float_as_int(t_crec[2 * k].x) >> 8) & 255) == 255) {
          float4 dk = dyn[k];
          dk.z -= sx; dk.w -= sy;
          dyn[k] = dk;
        }
      ox += sx; oy += sy;
    }
    KX_WAVE_SYNC();
  }

  // ── store ──
  for (int i = sub; i < nb; i += G) {
    int g = b0 + i;
    float4 p = prv[i], pp = pos[i], v = vel[i];
    d.px[g] = p.x + ox; d.py[g] = p.y + oy; d.ang[g] = p.z;
    d.vx[g] = v.x; d.vy[g] = v.y; d.av[g] = v.z;
    d.ppx[g] = pp.x + ox; d.ppy[g] = pp.y + oy; d.angP[g] = pp.z;
  }
  for (int k = sub; k < nc; k += G) {
    int g = c0 + k;
    int v = __float_as_int(t_crec[2 * k].x);
    float4 dk = dyn[k], lk = lamr[k];
    d.broken[g] = __float_as_int(lk.w);
    if (((v >> 8) & 255) == 255) { d.lbx[g] = dk.z + ox; d.lby[g] = dk.w + oy; }
    d.bf[g] = lk.z; d.target[g] = dk.y;
    if (has_law && (v >> 16) == KX_ANGLE) d.comp[g] = (float)L.comp;
    if ((v >> 16) == KX_ANGLE) { d.lam[g] = lk.x; d.lamx[g] = 0.0f; d.lamy[g] = 0.0f; }
    else { d.lamx[g] = lk.x; d.lamy[g] = lk.y; d.lam[g] = 0.0f; }
  }
  if (act && sub < T.narms) d.cd[cid * KX_MAX_ARMS + sub] = my_cd;
}

__global__ void kx_k_bump(int *counter, int delta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += delta;
}

int kx_bump_counter(int *counter, int delta, hipStream_t st) {
  hipLaunchKernelGGL(kx_k_bump, dim3(1), dim3(1), 0, st, counter, delta);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <int G, int WPB>
static int launch_gw(const KxDev &d, const KxLaw *law, int step0, const int *step_base, int nsteps, int ev_row0,
                     int with_forces, const KxPersistCfg &P, hipStream_t st) {
  constexpr int CPW = KX_WAVE / G;
  size_t lds = ((size_t)P.tmpl_words + P.bm_words + (size_t)WPB * CPW * P.cr_stride) * 4;
  if (lds > 160 * 1024) return -22;
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void *)kx_k_persist<G, WPB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    attr_set = true;
  }
  KxLaw L = law ? *law : KxLaw{};
  int blocks = (P.nwaves + WPB - 1) / WPB;
  hipLaunchKernelGGL((kx_k_persist<G, WPB>), dim3(blocks), dim3(KX_WAVE * WPB), lds, st, d, L, law ? 1 : 0, step0,
                     step_base, nsteps, ev_row0, with_forces, P);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <int G>
static int launch_g(int wpb, const KxDev &d, const KxLaw *law, int step0, const int *step_base, int nsteps,
                    int ev_row0, int with_forces, const KxPersistCfg &P, hipStream_t st) {
  switch (wpb) {
    case 1: return launch_gw<G, 1>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 2: return launch_gw<G, 2>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 4: return launch_gw<G, 4>(d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    default: return -22;
  }
}

int kx_persist_launch(int G, int wpb, const KxDev &d, const KxLaw *law, int step0, const int *step_base,
                      int nsteps, int ev_row0, int with_forces, const KxPersistCfg &P, hipStream_t st) {
  switch (G) {
    case 8: return launch_g<8>(wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 16: return launch_g<16>(wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 32: return launch_g<32>(wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    case 64: return launch_g<64>(wpb, d, law, step0, step_base, nsteps, ev_row0, with_forces, P, st);
    default: return -22;
  }
}
