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

// per-creature LDS layout (word offsets), arrays of NB bodies / NC constraints
struct KxCrLds {
  int px, py, ang, vx, vy, av, ppx, ppy, angP;
  int alpha, tgt, lbx, lby, bf, l0, l1, brk, ev, words;
  __host__ __device__ KxCrLds(int NB, int NC) {
    int o = 0;
    px = o; o += NB; py = o; o += NB; ang = o; o += NB; vx = o; o += NB; vy = o; o += NB; av = o; o += NB;
    ppx = o; o += NB; ppy = o; o += NB; angP = o; o += NB;
    alpha = o; o += NC; tgt = o; o += NC; lbx = o; o += NC; lby = o; o += NC; bf = o; o += NC;
    l0 = o; o += NC; l1 = o; o += NC; brk = o; o += NC;
    ev = o; o += 1;
    words = o;
  }
};

int kx_persist_cr_words(int NB, int NC) { return KxCrLds(NB, NC).words; }

template <int G, int WPB>
__global__ __launch_bounds__(KX_WAVE *WPB) void kx_k_persist(KxDev d, KxLaw L, int has_law, int step0,
                                                             const int *step_base, int nsteps, int ev_row0,
                                                             int with_forces, KxPersistCfg P) {
  constexpr int CPW = KX_WAVE / G;
  extern __shared__ uint32_t smem[];
  const int lane = threadIdx.x & (KX_WAVE - 1);
  const int q = lane / G, sub = lane % G;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / KX_WAVE);
  const int gw = blockIdx.x * WPB + w;

  // ── template -> LDS (the only workgroup barrier) ──
  const int tid = P.wave_tmpl[blockIdx.x * WPB];
  const KxTmpl T = P.tmpl[tid];
  for (int i = threadIdx.x; i < T.words; i += blockDim.x) smem[i] = P.pool[T.base + i];
  __syncthreads();
  if (gw >= P.nwaves) return;

  const int *t_cab = (const int *)(smem + T.o_cab);
  const float *t_lax = (const float *)(smem + T.o_lax), *t_lay = (const float *)(smem + T.o_lay);
  const float *t_rest = (const float *)(smem + T.o_rest), *t_mn = (const float *)(smem + T.o_mn);
  const float *t_mx = (const float *)(smem + T.o_mx);
  const float *t_im = (const float *)(smem + T.o_im), *t_ii = (const float *)(smem + T.o_ii);
  const int *t_bpo = (const int *)(smem + T.o_bpo);
  const float *t_plx = (const float *)(smem + T.o_plx), *t_ply = (const float *)(smem + T.o_ply);
  const float *t_pr = (const float *)(smem + T.o_pr);
  const int *t_sched = (const int *)(smem + T.o_sched), *t_rtype = (const int *)(smem + T.o_rtype);
  const int *t_coll = (const int *)(smem + T.o_coll), *t_ccnt = (const int *)(smem + T.o_ccnt);
  const int *t_arm_tip = (const int *)(smem + T.o_arm_tip), *t_arm_grip = (const int *)(smem + T.o_arm_grip);
  const int *t_mot_k = (const int *)(smem + T.o_mot_k), *t_mot_slot = (const int *)(smem + T.o_mot_slot);
  const int *t_grip_k = (const int *)(smem + T.o_grip_k), *t_grip_arm = (const int *)(smem + T.o_grip_arm);

  const int cid = P.wave_cr[gw * CPW + q];
  const bool act = cid >= 0;
  const KxCrLds Lo(P.NB, P.NC);
  float *S = (float *)(smem + P.tmpl_words + (size_t)(w * CPW + q) * P.cr_stride);
  float *px = S + Lo.px, *py = S + Lo.py, *ang = S + Lo.ang, *vx = S + Lo.vx, *vy = S + Lo.vy, *av = S + Lo.av;
  float *ppx = S + Lo.ppx, *ppy = S + Lo.ppy, *angP = S + Lo.angP;
  float *alpha = S + Lo.alpha, *tgt = S + Lo.tgt, *lbx = S + Lo.lbx, *lby = S + Lo.lby, *bf = S + Lo.bf;
  float *l0 = S + Lo.l0, *l1 = S + Lo.l1;
  int *brk = (int *)(S + Lo.brk);
  uint32_t *evw = (uint32_t *)(S + Lo.ev);
  const int nb = act ? T.nb : 0, nc = act ? T.nc : 0;
  const int b0 = act ? d.cr[cid].body_off : 0, c0 = act ? d.cr[cid].cons_off : 0;

  float ox = 0.0f, oy = 0.0f;
  if (act) { ox = rintf(d.px[b0]); oy = rintf(d.py[b0]); }

  // ── load the creature's dynamic state ──
  for (int i = sub; i < nb; i += G) {
    int g = b0 + i;
    px[i] = d.px[g] - ox; py[i] = d.py[g] - oy; ang[i] = d.ang[g];
    vx[i] = d.vx[g]; vy[i] = d.vy[g]; av[i] = d.av[g];
  }
  for (int k = sub; k < nc; k += G) {
    int g = c0 + k;
    bool world = ((t_cab[k] >> 8) & 255) == 255;
    alpha[k] = d.comp[g] * d.invH2; tgt[k] = d.target[g];
    lbx[k] = world ? d.lbx[g] - ox : d.lbx[g];
    lby[k] = world ? d.lby[g] - oy : d.lby[g];
    bf[k] = d.bf[g]; brk[k] = d.broken[g];
    l0[k] = 0.0f; l1[k] = 0.0f;
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
        tgt[k] = kx_law_target(L, (uint32_t)cid, (uint32_t)key, t_mot_slot[j]);
        alpha[k] = a_law;
      }
      // adhesion reflex (core.js:199-218), one lane per arm
      if (act && sub < T.narms) {
        double adh = kx_ctl_u(L.seed, (uint32_t)cid, (uint32_t)key, KX_SLOT_ADH + sub);
        int g = t_arm_grip[sub], t = t_arm_tip[sub];
        int bk = brk[g];
        float f_bf = bf[g], gx = lbx[g] + ox, gy = lby[g] + oy, lx = l0[g], ly = l1[g];
        uint32_t e = kx_reflex(L, hS, adh, my_cd, bk, f_bf, gx, gy, lx, ly, vx[t], vy[t],
                               (double)ox + (double)px[t], (double)oy + (double)py[t], ang[t], t_lax[g], t_lay[g]);
        brk[g] = bk; bf[g] = f_bf; lbx[g] = gx - ox; lby[g] = gy - oy; l0[g] = lx; l1[g] = ly;
        my_ev |= e << (3 * sub);
      }
      KX_WAVE_SYNC();
    }
    for (int s = 0; s < d.S; s++) {
      // resetLambda + integrate (xpbd.js:262-287)
      for (int k = sub; k < nc; k += G) { l0[k] = 0.0f; l1[k] = 0.0f; }
      const bool forces = with_forces && step == 0 && s == 0;
      for (int i = sub; i < nb; i += G) {
        float fx = 0.0f, fy = 0.0f, tq = 0.0f;
        if (forces) {
          int g = b0 + i;
          fx = d.fx[g]; fy = d.fy[g]; tq = d.tq[g];
          d.fx[g] = 0.0f; d.fy[g] = 0.0f; d.tq[g] = 0.0f;
        }
        float a_vx = vx[i], a_vy = vy[i], a_av = av[i], q_x, q_y, q_a;
        kx_integrate(d, a_vx, a_vy, a_av, px[i], py[i], ang[i], t_im[i], t_ii[i], fx, fy, tq, q_x, q_y, q_a);
        vx[i] = a_vx; vy[i] = a_vy; av[i] = a_av; ppx[i] = q_x; ppy[i] = q_y; angP[i] = q_a;
      }
      KX_WAVE_SYNC();
      for (int it = 0; it < d.I; it++) {
        // constraint rounds (type-homogeneous, wave-uniform branch)
        for (int r = 0; r < T.R; r++) {
          const int k = act ? t_sched[r * G + sub] : -1;
          if (t_rtype[r]) {
            if (k >= 0 && !brk[k]) {
              int v = t_cab[k];
              int A = v & 255, B = (v >> 8) & 255;
              float lam = l0[k];
              kx_solve_angle(angP, t_ii, A, B == 255 ? -1 : B, t_rest[k], tgt[k], t_mn[k], t_mx[k], alpha[k], lam);
              l0[k] = lam;
            }
          } else {
            if (k >= 0 && !brk[k]) {
              int v = t_cab[k];
              int A = v & 255, B = (v >> 8) & 255;
              float lx = l0[k], ly = l1[k];
              kx_solve_pin(ppx, ppy, angP, t_im, t_ii, A, B == 255 ? -1 : B, t_lax[k], t_lay[k], lbx[k], lby[k],
                           alpha[k], lx, ly);
              l0[k] = lx; l1[k] = ly;
            }
          }
          KX_WAVE_SYNC();
        }
        // collisions: this lane's bodies, each body's points in order
        const int nbl = act ? t_ccnt[sub] : 0;
        for (int j = 0; j < nbl; j++) {
          int i = t_coll[sub * T.maxbl + j];
          if (t_im[i] == 0.0f) continue;
          int bp = t_bpo[i];
          int o = bp & 0xffff, cnt = bp >> 16;
          float x = ppx[i], y = ppy[i];
          float sa, ca;
          sincosf(angP[i], &sa, &ca);
          for (int pq = 0; pq < cnt; pq++) {
            float lx = t_plx[o + pq], ly = t_ply[o + pq];
            float wx = x + (lx * ca - ly * sa);
            float wy = y + (lx * sa + ly * ca);
            KxHit hh;
            if (!kx_sdf_grid(d, wx, wy, ox, oy, hh)) continue;
            float pen = t_pr[o + pq] - hh.d;
            if (pen > KX_SLACK_POS) { x += hh.nx * pen; y += hh.ny * pen; }
          }
          ppx[i] = x; ppy[i] = y;
        }
        KX_WAVE_SYNC();
      }
      for (int i = sub; i < nb; i += G) {
        float a_px = px[i], a_py = py[i], a_ang = ang[i], a_vx, a_vy, a_av, a_angP = angP[i];
        kx_finalize(d, a_px, a_py, a_ang, a_vx, a_vy, a_av, ppx[i], ppy[i], a_angP);
        px[i] = a_px; py[i] = a_py; ang[i] = a_ang; angP[i] = a_angP;
        vx[i] = a_vx; vy[i] = a_vy; av[i] = a_av;
      }
      KX_WAVE_SYNC();
    }
    // grip break (xpbd.js:496-504 / solver.wgsl:433-442)
    for (int j = sub; j < (act ? T.ngrip : 0); j += G) {
      int k = t_grip_k[j];
      if (brk[k]) continue;
      float lx = l0[k], ly = l1[k];
      float f = sqrtf(lx * lx + ly * ly) * d.invH;
      if (f > bf[k]) {
        brk[k] = 1; l0[k] = 0.0f; l1[k] = 0.0f;
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
    float sx = act ? rintf(px[0]) : 0.0f, sy = act ? rintf(py[0]) : 0.0f;
    KX_WAVE_SYNC();
    if (fabsf(sx) > 32.0f || fabsf(sy) > 32.0f) {
      for (int i = sub; i < nb; i += G) { px[i] -= sx; py[i] -= sy; ppx[i] -= sx; ppy[i] -= sy; }
      for (int k = sub; k < nc; k += G)
        if (((t_cab[k] >> 8) & 255) == 255) { lbx[k] -= sx; lby[k] -= sy; }
      ox += sx; oy += sy;
    }
    KX_WAVE_SYNC();
  }

  // ── store ──
  for (int i = sub; i < nb; i += G) {
    int g = b0 + i;
    d.px[g] = px[i] + ox; d.py[g] = py[i] + oy; d.ang[g] = ang[i];
    d.vx[g] = vx[i]; d.vy[g] = vy[i]; d.av[g] = av[i];
    d.ppx[g] = ppx[i] + ox; d.ppy[g] = ppy[i] + oy; d.angP[g] = angP[i];
  }
  for (int k = sub; k < nc; k += G) {
    int g = c0 + k;
    int v = t_cab[k];
    d.broken[g] = brk[k];
    if (((v >> 8) & 255) == 255) { d.lbx[g] = lbx[k] + ox; d.lby[g] = lby[k] + oy; }
    d.bf[g] = bf[k]; d.target[g] = tgt[k];
    if (has_law && (v >> 16) == KX_ANGLE) d.comp[g] = (float)L.comp;
    if ((v >> 16) == KX_ANGLE) { d.lam[g] = l0[k]; d.lamx[g] = 0.0f; d.lamy[g] = 0.0f; }
    else { d.lamx[g] = l0[k]; d.lamy[g] = l1[k]; d.lam[g] = 0.0f; }
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
  size_t lds = ((size_t)P.tmpl_words + (size_t)WPB * CPW * P.cr_stride) * 4;
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
