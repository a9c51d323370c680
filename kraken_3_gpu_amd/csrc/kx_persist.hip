// Persistent batched path: one wavefront owns one creature and keeps its
// bodies, constraints and collision points in LDS for every substep and
// iteration of a multi-step launch. Colour passes (the sequential
// per-colour dispatches of gpuworld.js:353-366) become in-wave passes: a
// wave's LDS traffic is processed in issue order, so a compiler fence between
// passes is all the ordering they need; no workgroup barrier is used.
// State is read from and written to HBM (SoA) only at launch boundaries.
//
// Positions live in creature-local coordinates relative to an integer
// origin (re-centred on the first body between steps); since every global
// binary32 coordinate is a multiple of its own ulp, the shift to and from the
// origin is exact and the solver works with ~8x finer position ulps.
#include "kx_device.h"
#include "../../include/kxpbd.h"

#define KX_WAVE_SYNC()                                     \
  do {                                                     \
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); \
    __builtin_amdgcn_wave_barrier();                       \
  } while (0)

// LDS sub-array offsets (in 4-byte words) for launch maxima NB, NC, NP
struct KxLds {
  int px, py, ang, vx, vy, av, ppx, ppy, angP, im, ii, bpo;
  int cab, brk, lax, lay, lbx, lby, comp, rest, mn, mx, tgt, bf, l0, l1, slot;
  int plx, ply, pr;
  int ev, stride;
  __device__ __host__ KxLds(int NB, int NC, int NP) {
    int o = 0;
    px = o; o += NB; py = o; o += NB; ang = o; o += NB; vx = o; o += NB; vy = o; o += NB; av = o; o += NB;
    ppx = o; o += NB; ppy = o; o += NB; angP = o; o += NB; im = o; o += NB; ii = o; o += NB; bpo = o; o += NB;
    cab = o; o += NC; brk = o; o += NC; lax = o; o += NC; lay = o; o += NC; lbx = o; o += NC; lby = o; o += NC;
    comp = o; o += NC; rest = o; o += NC; mn = o; o += NC; mx = o; o += NC; tgt = o; o += NC; bf = o; o += NC;
    l0 = o; o += NC; l1 = o; o += NC; slot = o; o += NC;
    plx = o; o += NP; ply = o; o += NP; pr = o; o += NP;
    ev = o; o += 1;
    stride = (o + 3) & ~3;
  }
};

int kx_persist_lds_words(int NB, int NC, int NP) { return KxLds(NB, NC, NP).stride; }

template <int WPB>
__global__ __launch_bounds__(KX_WAVE *WPB) void kx_k_persist(KxDev d, KxLaw L, int has_law, int step0,
                                                             const int *step_base, int nsteps, int ev_row0,
                                                             int with_forces, int NB, int NC, int NP) {
  extern __shared__ float smem[];
  const int lane = threadIdx.x & (KX_WAVE - 1);
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / KX_WAVE);
  const int cid = __builtin_amdgcn_readfirstlane(blockIdx.x * WPB + w);
  if (cid >= d.nCr) return;                       // whole wave leaves; no block barriers below
  const KxLds Lo(NB, NC, NP);
  float *S = smem + (size_t)w * Lo.stride;
  int *Si = reinterpret_cast<int *>(S);
  const KxCreature cr = d.cr[cid];
  const int b0 = cr.body_off, nb = cr.nb, c0 = cr.cons_off, nc = cr.nc, p0 = cr.pt_off, np = cr.np;

  float *px = S + Lo.px, *py = S + Lo.py, *ang = S + Lo.ang, *vx = S + Lo.vx, *vy = S + Lo.vy, *av = S + Lo.av;
  float *ppx = S + Lo.ppx, *ppy = S + Lo.ppy, *angP = S + Lo.angP, *im = S + Lo.im, *ii = S + Lo.ii;
  int *bpo = Si + Lo.bpo;
  int *cab = Si + Lo.cab, *brk = Si + Lo.brk, *slot = Si + Lo.slot;
  float *lax = S + Lo.lax, *lay = S + Lo.lay, *lbx = S + Lo.lbx, *lby = S + Lo.lby, *comp = S + Lo.comp;
  float *rest = S + Lo.rest, *mn = S + Lo.mn, *mx = S + Lo.mx, *tgt = S + Lo.tgt, *bf = S + Lo.bf;
  float *l0 = S + Lo.l0, *l1 = S + Lo.l1;
  float *plx = S + Lo.plx, *ply = S + Lo.ply, *pr = S + Lo.pr;
  uint32_t *evw = reinterpret_cast<uint32_t *>(Si + Lo.ev);

  // integer origin: the first body's position rounded
  float ox = rintf(d.px[b0]), oy = rintf(d.py[b0]);

  // ── load ──
  for (int i = lane; i < nb; i += KX_WAVE) {
    int g = b0 + i;
    px[i] = d.px[g] - ox; py[i] = d.py[g] - oy; ang[i] = d.ang[g];
    vx[i] = d.vx[g]; vy[i] = d.vy[g]; av[i] = d.av[g];
    im[i] = d.im[g]; ii[i] = d.ii[g];
    bpo[i] = (d.bpoff[g] - p0) | (d.bpcnt[g] << 16);
  }
  for (int k = lane; k < nc; k += KX_WAVE) {
    int g = c0 + k;
    int a = d.ca[g] - b0, bg = d.cb[g];
    int b = bg < 0 ? 255 : bg - b0;
    int t = d.ctype[g];
    cab[k] = a | (b << 8) | (t << 16);
    brk[k] = d.broken[g];
    lax[k] = d.lax[g]; lay[k] = d.lay[g];
    float bx = d.lbx[g], by = d.lby[g];
    if (bg < 0) { bx -= ox; by -= oy; }
    lbx[k] = bx; lby[k] = by;
    comp[k] = d.comp[g]; rest[k] = d.rest[g]; mn[k] = d.cmin[g]; mx[k] = d.cmax[g];
    tgt[k] = d.target[g]; bf[k] = d.bf[g];
    l0[k] = 0.0f; l1[k] = 0.0f;
    int s = d.cslot[g];
    if (s < 0 && d.carm[g] >= 0) s = -2 - d.carm[g];
    slot[k] = s;
  }
  for (int p = lane; p < np; p += KX_WAVE) {
    plx[p] = d.plx[p0 + p]; ply[p] = d.ply[p0 + p]; pr[p] = d.pr[p0 + p];
  }
  // arms: one lane per arm keeps tip, grip and the f64 cooldown in registers
  int my_tip = 0, my_grip = 0;
  double my_cd = 0.0;
  if (lane < cr.narms) {
    my_tip = d.arm_tip[cr.arm_off + lane] - b0;
    my_grip = d.arm_grip[cr.arm_off + lane] - c0;
    my_cd = d.cd[cid * KX_MAX_ARMS + lane];
  }
  uint16_t coff[KX_MAX_COLORS + 1];
#pragma unroll
  for (int c = 0; c <= KX_MAX_COLORS; c++) coff[c] = cr.coff[c];
  const int ncol = cr.ncol;
  const double hS = d.hd * d.S;
  KX_WAVE_SYNC();

  for (int step = 0; step < nsteps; step++) {
    const int key = (step_base ? step_base[0] : step0) + step;
    if (lane == 0) evw[0] = 0u;
    uint32_t my_ev = 0u;
    if (has_law) {
      // motor targets + compliance (core.js:190-198)
      for (int k = lane; k < nc; k += KX_WAVE) {
        int s = slot[k];
        if (s >= 0) { tgt[k] = kx_law_target(L, (uint32_t)cid, (uint32_t)key, s); comp[k] = (float)L.comp; }
      }
      KX_WAVE_SYNC();
      // adhesion reflex (core.js:199-218)
      if (lane < cr.narms) {
        double adh = kx_ctl_u(L.seed, (uint32_t)cid, (uint32_t)key, KX_SLOT_ADH + lane);
        int g = my_grip, t = my_tip;
        int bk = brk[g];
        float f_bf = bf[g], gx = lbx[g] + ox, gy = lby[g] + oy, lx = l0[g], ly = l1[g];
        uint32_t e = kx_reflex(L, hS, adh, my_cd, bk, f_bf, gx, gy, lx, ly, vx[t], vy[t],
                               (double)ox + (double)px[t], (double)oy + (double)py[t], ang[t], lax[g], lay[g]);
        brk[g] = bk; bf[g] = f_bf; lbx[g] = gx - ox; lby[g] = gy - oy; l0[g] = lx; l1[g] = ly;
        my_ev |= e << (3 * lane);
      }
      KX_WAVE_SYNC();
    }
    for (int s = 0; s < d.S; s++) {
      // resetLambda + integrate
      for (int k = lane; k < nc; k += KX_WAVE) { l0[k] = 0.0f; l1[k] = 0.0f; }
      const bool forces = with_forces && step == 0 && s == 0;
      for (int i = lane; i < nb; i += KX_WAVE) {
        float fx = 0.0f, fy = 0.0f, tq = 0.0f;
        if (forces) {
          int g = b0 + i;
          fx = d.fx[g]; fy = d.fy[g]; tq = d.tq[g];
          d.fx[g] = 0.0f; d.fy[g] = 0.0f; d.tq[g] = 0.0f;
        }
        float a_vx = vx[i], a_vy = vy[i], a_av = av[i], q_x, q_y, q_a;
        kx_integrate(d, a_vx, a_vy, a_av, px[i], py[i], ang[i], im[i], ii[i], fx, fy, tq, q_x, q_y, q_a);
        vx[i] = a_vx; vy[i] = a_vy; av[i] = a_av; ppx[i] = q_x; ppy[i] = q_y; angP[i] = q_a;
      }
      KX_WAVE_SYNC();
      for (int it = 0; it < d.I; it++) {
        for (int c = 0; c < ncol; c++) {
          const int j0 = coff[c], cnt = coff[c + 1] - j0;
          for (int j = lane; j < cnt; j += KX_WAVE) {
            int k = j0 + j;
            if (brk[k]) continue;
            int v = cab[k];
            int A = v & 255, B = (v >> 8) & 255, t = v >> 16;
            if (B == 255) B = -1;
            float alpha = comp[k] * d.invH2;
            if (t == KX_ANGLE) {
              float lam = l0[k];
              kx_solve_angle(angP, ii, A, B, rest[k], tgt[k], mn[k], mx[k], alpha, lam);
              l0[k] = lam;
            } else {
              float lx = l0[k], ly = l1[k];
              kx_solve_pin(ppx, ppy, angP, im, ii, A, B, lax[k], lay[k], lbx[k], lby[k], alpha, lx, ly);
              l0[k] = lx; l1[k] = ly;
            }
          }
          KX_WAVE_SYNC();
        }
        // collisions: one lane per body, its points in order (cumulative)
        for (int i = lane; i < nb; i += KX_WAVE) {
          int bp = bpo[i];
          int o = bp & 0xffff, cnt = bp >> 16;
          if (cnt == 0 || im[i] == 0.0f) continue;
          float x = ppx[i], y = ppy[i];
          float sa, ca;
          sincosf(angP[i], &sa, &ca);
          for (int q = 0; q < cnt; q++) {
            float lx = plx[o + q], ly = ply[o + q];
            float wx = x + (lx * ca - ly * sa);
            float wy = y + (lx * sa + ly * ca);
            KxHit h;
            if (!kx_sdf_grid(d, wx, wy, ox, oy, h)) continue;
            float pen = pr[o + q] - h.d;
            if (pen > KX_SLACK_POS) { x += h.nx * pen; y += h.ny * pen; }
          }
          ppx[i] = x; ppy[i] = y;
        }
        KX_WAVE_SYNC();
      }
      for (int i = lane; i < nb; i += KX_WAVE) {
        float a_px = px[i], a_py = py[i], a_ang = ang[i], a_vx, a_vy, a_av, a_angP = angP[i];
        kx_finalize(d, a_px, a_py, a_ang, a_vx, a_vy, a_av, ppx[i], ppy[i], a_angP);
        px[i] = a_px; py[i] = a_py; ang[i] = a_ang; angP[i] = a_angP;
        vx[i] = a_vx; vy[i] = a_vy; av[i] = a_av;
      }
      KX_WAVE_SYNC();
    }
    // grip break (xpbd.js:496-504 / solver.wgsl:433-442)
    for (int k = lane; k < nc; k += KX_WAVE) {
      if ((cab[k] >> 16) != KX_GRIP || brk[k]) continue;
      float lx = l0[k], ly = l1[k];
      float f = sqrtf(lx * lx + ly * ly) * d.invH;
      if (f > bf[k]) {
        brk[k] = 1; l0[k] = 0.0f; l1[k] = 0.0f;
        int sl = slot[k];
        if (sl <= -2) my_ev |= 4u << (3 * (-2 - sl));
      }
    }
    if (d.events) {
      if (my_ev) atomicOr(evw, my_ev);
      KX_WAVE_SYNC();
      if (lane == 0) d.events[(size_t)(ev_row0 + step) * d.nCr + cid] = evw[0];
    }
    // re-centre the origin on the first body when it drifted
    float sx = rintf(px[0]), sy = rintf(py[0]);
    if (fabsf(sx) > 32.0f || fabsf(sy) > 32.0f) {
      KX_WAVE_SYNC();
      for (int i = lane; i < nb; i += KX_WAVE) { px[i] -= sx; py[i] -= sy; ppx[i] -= sx; ppy[i] -= sy; }
      for (int k = lane; k < nc; k += KX_WAVE) {
        if (((cab[k] >> 8) & 255) == 255) { lbx[k] -= sx; lby[k] -= sy; }
      }
      ox += sx; oy += sy;
      KX_WAVE_SYNC();
    }
  }

  // ── store ──
  for (int i = lane; i < nb; i += KX_WAVE) {
    int g = b0 + i;
    d.px[g] = px[i] + ox; d.py[g] = py[i] + oy; d.ang[g] = ang[i];
    d.vx[g] = vx[i]; d.vy[g] = vy[i]; d.av[g] = av[i];
    d.ppx[g] = ppx[i] + ox; d.ppy[g] = ppy[i] + oy; d.angP[g] = angP[i];
  }
  for (int k = lane; k < nc; k += KX_WAVE) {
    int g = c0 + k;
    d.broken[g] = brk[k];
    if (((cab[k] >> 8) & 255) == 255) { d.lbx[g] = lbx[k] + ox; d.lby[g] = lby[k] + oy; }
    d.bf[g] = bf[k]; d.target[g] = tgt[k]; d.comp[g] = comp[k];
    int t = cab[k] >> 16;
    if (t == KX_ANGLE) { d.lam[g] = l0[k]; d.lamx[g] = 0.0f; d.lamy[g] = 0.0f; }
    else { d.lamx[g] = l0[k]; d.lamy[g] = l1[k]; d.lam[g] = 0.0f; }
  }
  if (lane < cr.narms) d.cd[cid * KX_MAX_ARMS + lane] = my_cd;
}

__global__ void kx_k_bump(int *counter, int delta) {
  if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] += delta;
}

int kx_bump_counter(int *counter, int delta, hipStream_t st) {
  hipLaunchKernelGGL(kx_k_bump, dim3(1), dim3(1), 0, st, counter, delta);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

template <int WPB>
static int launch_wpb(const KxDev &d, const KxLaw *law, int step0, const int *step_base, int nsteps, int ev_row0,
                      int with_forces, int NB, int NC, int NP, hipStream_t st) {
  KxLds Lo(NB, NC, NP);
  size_t lds = (size_t)Lo.stride * 4 * WPB;
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void *)kx_k_persist<WPB>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  KxLaw L = law ? *law : KxLaw{};
  int blocks = (d.nCr + WPB - 1) / WPB;
  hipLaunchKernelGGL(kx_k_persist<WPB>, dim3(blocks), dim3(KX_WAVE * WPB), lds, st, d, L, law ? 1 : 0, step0,
                     step_base, nsteps, ev_row0, with_forces, NB, NC, NP);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

int kx_persist_launch(const KxDev &d, const KxLaw *law, int step0, const int *step_base_dev, int nsteps,
                      int ev_row0, int waves_per_block, int with_forces, int maxb, int maxc, int maxp,
                      hipStream_t st) {
  switch (waves_per_block) {
    case 1: return launch_wpb<1>(d, law, step0, step_base_dev, nsteps, ev_row0, with_forces, maxb, maxc, maxp, st);
    case 2: return launch_wpb<2>(d, law, step0, step_base_dev, nsteps, ev_row0, with_forces, maxb, maxc, maxp, st);
    case 4: return launch_wpb<4>(d, law, step0, step_base_dev, nsteps, ev_row0, with_forces, maxb, maxc, maxp, st);
    case 8: return launch_wpb<8>(d, law, step0, step_base_dev, nsteps, ev_row0, with_forces, maxb, maxc, maxp, st);
    default: return -22;
  }
}
