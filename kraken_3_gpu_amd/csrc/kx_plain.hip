// Plain per-colour-launch path: the launch structure of the reference's
// WebGPU solver (src/gpu/gpuworld.js:344-382 substep/step driving
// src/gpu/solver.wgsl), rebuilt as HIP kernels over SoA buffers.
// Per step: [control law] + to-local + S x (reset, integrate,
// I x (one launch per non-empty colour, collisions), finalize) + grip break
// + to-global. Between to-local and to-global the position arrays hold
// creature-local coordinates (integer origin per creature, exact shifts).
#include "kx_device.h"

#define PLAIN_BLOCK 256

static inline int nblk(int n) { return (n + PLAIN_BLOCK - 1) / PLAIN_BLOCK; }

// Control law + reflex, one thread per arm (core.js:183-220).
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_arms(KxDev d, KxLaw L, int step_key, const int *step_base,
                                                         int ev_row, int tape_row) {
  int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= d.nArms) return;
  if (step_base) step_key += step_base[0];
  // creature of this arm: binary search over arm_off
  int lo = 0, hi = d.nCr - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (d.cr[mid].arm_off <= a) lo = mid; else hi = mid - 1;
  }
  int c = lo, arm = a - d.cr[c].arm_off;
  int tip = d.arm_tip[a], g = d.arm_grip[a];
  const uint32_t ck = (uint32_t)c + L.cbase;
  double adh = (L.flags & KX_LAW_TAPE) ? (double)kx_tape_rec(d, tape_row, c)[3 + 3 * arm]
                                       : kx_ctl_u(L.seed, ck, (uint32_t)step_key, KX_SLOT_ADH + arm);
  double gu = (L.flags & KX_LAW_FIXED_GRIP) ? kx_ctl_u(L.seed, ck, KX_GRIP_DRAW_KEY, KX_SLOT_ADH + arm) : 0.0;
  double &cd = d.cd[c * KX_MAX_ARMS + arm];
  int broken = d.broken[g];
  float bf = d.bf[g], lbx = d.lbx[g], lby = d.lby[g], lx = d.lamx[g], ly = d.lamy[g];
  uint32_t ev = kx_reflex(L, d.hd * d.S, adh, gu, cd, broken, bf, lbx, lby, lx, ly, d.vx[tip], d.vy[tip],
                          (double)d.px[tip], (double)d.py[tip], d.ang[tip], d.lax[g], d.lay[g]);
  d.broken[g] = broken; d.bf[g] = bf; d.lbx[g] = lbx; d.lby[g] = lby; d.lamx[g] = lx; d.lamy[g] = ly;
  if (ev && d.events) atomicOr(&d.events[(size_t)ev_row * d.nCr + c], ev << (3 * arm));
}

// Motor targets + compliance under the law or the tape, one thread per constraint.
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_motors(KxDev d, KxLaw L, int step_key, const int *step_base,
                                                           int tape_row) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.nC) return;
  if (step_base) step_key += step_base[0];
  int slot = d.cslot[k];
  if (slot < 0) return;
  if (L.flags & KX_LAW_TAPE) {
    const float *rec = kx_tape_rec(d, tape_row, d.ccr[k]);
    d.target[k] = kx_tape_target(L, rec, slot);
    d.comp[k] = (float)kx_tape_comp(L, rec[0]);
    return;
  }
  d.target[k] = kx_law_target(L, (uint32_t)d.ccr[k] + L.cbase, (uint32_t)step_key, slot);
  d.comp[k] = (float)L.comp;
}

// Integer origin per creature: its first body's position, rounded (0 when
// not finite).
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_origin(KxDev d) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= d.nCr) return;
  int b = d.cr[c].body_off;
  d.org[2 * c] = kx_origin(d.px[b]);
  d.org[2 * c + 1] = kx_origin(d.py[b]);
}

// Shift bodies and world anchors by -origin (sign=-1) or +origin (sign=+1).
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_shift(KxDev d, float sign) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < d.nB) {
    int c = d.bcr[i];
    float ox = sign * d.org[2 * c], oy = sign * d.org[2 * c + 1];
    d.px[i] += ox; d.py[i] += oy; d.ppx[i] += ox; d.ppy[i] += oy;
  }
  if (i < d.nC && d.cb[i] < 0) {
    int c = d.ccr[i];
    d.lbx[i] += sign * d.org[2 * c];
    d.lby[i] += sign * d.org[2 * c + 1];
  }
}

// resetLambda (solver.wgsl:156-163)
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_reset(KxDev d) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.nC) return;
  d.lam[k] = 0.0f; d.lamx[k] = 0.0f; d.lamy[k] = 0.0f;
}

// integrate (solver.wgsl:166-205); forces consumed on the first substep
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_integrate(KxDev d, int with_forces) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.nB) return;
  float vx = d.vx[i], vy = d.vy[i], av = d.av[i], ppx, ppy, angP;
  float fx = 0.0f, fy = 0.0f, tq = 0.0f;
  if (with_forces) { fx = d.fx[i]; fy = d.fy[i]; tq = d.tq[i]; d.fx[i] = 0.0f; d.fy[i] = 0.0f; d.tq[i] = 0.0f; }
  kx_integrate(d, vx, vy, av, d.px[i], d.py[i], d.ang[i], d.im[i], d.ii[i], fx, fy, tq, ppx, ppy, angP);
  d.vx[i] = vx; d.vy[i] = vy; d.av[i] = av;
  d.ppx[i] = ppx; d.ppy[i] = ppy; d.angP[i] = angP;
}

// solveConstraints for one colour (solver.wgsl:208-350)
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_colour(KxDev d, int off, int cnt) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= cnt) return;
  int k = d.col_list[off + j];
  if (d.broken[k]) return;
  int A = d.ca[k], B = d.cb[k];
  float alpha = d.comp[k] * d.invH2;
  if (d.ctype[k] == 1) {
    float lam = d.lam[k];
    kx_solve_angle(d.angP, d.ii, A, B, d.rest[k], d.target[k], d.cmin[k], d.cmax[k], alpha, lam);
    d.lam[k] = lam;
  } else {
    float lx = d.lamx[k], ly = d.lamy[k];
    kx_solve_pin(d.ppx, d.ppy, d.angP, d.im, d.ii, A, B, d.lax[k], d.lay[k], d.lbx[k], d.lby[k], alpha, lx, ly);
    d.lamx[k] = lx; d.lamy[k] = ly;
  }
}

// solveCollisions, one thread per body, its points in order (solver.wgsl:373-402)
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_collide(KxDev d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.nB) return;
  if (d.im[i] == 0.0f) return;
  int cnt = d.bpcnt[i];
  if (cnt == 0) return;
  int o = d.bpoff[i];
  int c = d.bcr[i];
  float ox = d.org[2 * c], oy = d.org[2 * c + 1];
  float x = d.ppx[i], y = d.ppy[i];
  float sa, ca;
  kx_sincos(d.angP[i], &sa, &ca);
  for (int q = 0; q < cnt; q++) {
    float wx, wy;
    kx_world_point(x, y, ca, sa, d.plx[o + q], d.ply[o + q], wx, wy);
    KxHit h;
    bool ok = d.use_grid ? kx_sdf_grid(d, wx, wy, ox, oy, h) : kx_sdf_linear(d, wx, wy, ox, oy, h);
    if (!ok) continue;
    float pen = d.pr[o + q] - h.d;
    if (pen > KX_SLACK_POS) {
      float nx, ny;
      kx_hit_normal(h, nx, ny);
      x += nx * pen; y += ny * pen;
    }
  }
  d.ppx[i] = x; d.ppy[i] = y;
}

// finalize (solver.wgsl:405-430)
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_finalize(KxDev d) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= d.nB) return;
  float px = d.px[i], py = d.py[i], ang = d.ang[i], vx, vy, av, angP = d.angP[i];
  kx_finalize(d, px, py, ang, vx, vy, av, d.ppx[i], d.ppy[i], angP);
  d.px[i] = px; d.py[i] = py; d.ang[i] = ang; d.angP[i] = angP;
  d.vx[i] = vx; d.vy[i] = vy; d.av[i] = av;
}

// gripBreak (solver.wgsl:433-442): F = |lambda|/h over the last substep
__global__ __launch_bounds__(PLAIN_BLOCK) void kx_k_grip_break(KxDev d, int ev_row) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= d.nC) return;
  if (d.ctype[k] != 2 || d.broken[k]) return;
  float lx = d.lamx[k], ly = d.lamy[k];
  float f = sqrtf(lx * lx + ly * ly) * d.invH;
  if (f > d.bf[k]) {
    d.broken[k] = 1; d.lamx[k] = 0.0f; d.lamy[k] = 0.0f;
    int arm = d.carm[k];
    if (d.events && arm >= 0) atomicOr(&d.events[(size_t)ev_row * d.nCr + d.ccr[k]], 4u << (3 * arm));
  }
}

int kx_plain_step(const KxDev &d, const KxLaw *law, int step_key, const int *step_base, int with_forces,
                  int ev_row, int tape_row, hipStream_t st, int *launches) {
  int n = 0;
  if (law) {
    hipLaunchKernelGGL(kx_k_motors, dim3(nblk(d.nC)), dim3(PLAIN_BLOCK), 0, st, d, *law, step_key, step_base,
                       tape_row); n++;
    if (d.nArms > 0) {
      hipLaunchKernelGGL(kx_k_arms, dim3(nblk(d.nArms)), dim3(PLAIN_BLOCK), 0, st, d, *law, step_key, step_base,
                         ev_row, tape_row); n++;
    }
  }
  int nbc = d.nB > d.nC ? d.nB : d.nC;
  hipLaunchKernelGGL(kx_k_origin, dim3(nblk(d.nCr)), dim3(PLAIN_BLOCK), 0, st, d); n++;
  hipLaunchKernelGGL(kx_k_shift, dim3(nblk(nbc)), dim3(PLAIN_BLOCK), 0, st, d, -1.0f); n++;
  for (int s = 0; s < d.S; s++) {
    hipLaunchKernelGGL(kx_k_reset, dim3(nblk(d.nC)), dim3(PLAIN_BLOCK), 0, st, d); n++;
    hipLaunchKernelGGL(kx_k_integrate, dim3(nblk(d.nB)), dim3(PLAIN_BLOCK), 0, st, d, with_forces && s == 0); n++;
    for (int it = 0; it < d.I; it++) {
      for (int c = 0; c < d.ncol; c++) {
        hipLaunchKernelGGL(kx_k_colour, dim3(nblk(d.col_cnt[c])), dim3(PLAIN_BLOCK), 0, st, d, d.col_off[c], d.col_cnt[c]); n++;
      }
      hipLaunchKernelGGL(kx_k_collide, dim3(nblk(d.nB)), dim3(PLAIN_BLOCK), 0, st, d); n++;
    }
    hipLaunchKernelGGL(kx_k_finalize, dim3(nblk(d.nB)), dim3(PLAIN_BLOCK), 0, st, d); n++;
  }
  hipLaunchKernelGGL(kx_k_grip_break, dim3(nblk(d.nC)), dim3(PLAIN_BLOCK), 0, st, d, ev_row); n++;
  hipLaunchKernelGGL(kx_k_shift, dim3(nblk(nbc)), dim3(PLAIN_BLOCK), 0, st, d, 1.0f); n++;
  if (launches) *launches += n;
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

// n substeps without the grip break (gpuworld.js:344-372 substep), in the
// creature-local frame like a full step; forces consumed by the first.
int kx_plain_substeps(const KxDev &d, int n, int with_forces, hipStream_t st, int *launches) {
  int k = 0;
  const int nbc = d.nB > d.nC ? d.nB : d.nC;
  hipLaunchKernelGGL(kx_k_origin, dim3(nblk(d.nCr)), dim3(PLAIN_BLOCK), 0, st, d); k++;
  hipLaunchKernelGGL(kx_k_shift, dim3(nblk(nbc)), dim3(PLAIN_BLOCK), 0, st, d, -1.0f); k++;
  for (int s = 0; s < n; s++) {
    hipLaunchKernelGGL(kx_k_reset, dim3(nblk(d.nC)), dim3(PLAIN_BLOCK), 0, st, d); k++;
    hipLaunchKernelGGL(kx_k_integrate, dim3(nblk(d.nB)), dim3(PLAIN_BLOCK), 0, st, d, with_forces && s == 0); k++;
    for (int it = 0; it < d.I; it++) {
      for (int c = 0; c < d.ncol; c++) {
        hipLaunchKernelGGL(kx_k_colour, dim3(nblk(d.col_cnt[c])), dim3(PLAIN_BLOCK), 0, st, d, d.col_off[c], d.col_cnt[c]); k++;
      }
      hipLaunchKernelGGL(kx_k_collide, dim3(nblk(d.nB)), dim3(PLAIN_BLOCK), 0, st, d); k++;
    }
    hipLaunchKernelGGL(kx_k_finalize, dim3(nblk(d.nB)), dim3(PLAIN_BLOCK), 0, st, d); k++;
  }
  hipLaunchKernelGGL(kx_k_shift, dim3(nblk(nbc)), dim3(PLAIN_BLOCK), 0, st, d, 1.0f); k++;
  if (launches) *launches += k;
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

// ── state readout ───────────────────────────────────────────────────
// readProprio (src/sim/core.js:129-175; tipPosition, armCurl, armExtension
// src/body/kraken.js:150-183) of every creature, in binary64 from the
// binary32 state, then its mantle state; one thread per arm (eight arm slots
// per creature, eight creatures per block), the arm sums taken in arm order by
// the creature's first thread. Math.min /
// Math.max give NaN for a NaN argument, unlike fmin / fmax, so the clamps are
// written as comparisons, as the reference's own `cl` is.
__device__ __forceinline__ double kx_js_min(double a, double b) { return (a != a || b != b) ? a + b : (a < b ? a : b); }
__device__ __forceinline__ double kx_js_max(double a, double b) { return (a != a || b != b) ? a + b : (a > b ? a : b); }
__device__ __forceinline__ double kx_cl(double v, double m) { return v > m ? m : (v < -m ? -m : v); }

constexpr int KX_RS_CR = 64 / KX_MAX_ARMS;   // creatures per 64-thread block, one thread per arm slot

__global__ __launch_bounds__(64) void kx_k_read_state(KxDev d, float *out, int stride) {
  __shared__ double s_ext[KX_RS_CR][KX_MAX_ARMS], s_curl[KX_RS_CR][KX_MAX_ARMS];
  __shared__ int s_brk[KX_RS_CR][KX_MAX_ARMS];
  const int lc = threadIdx.x / KX_MAX_ARMS, a = threadIdx.x % KX_MAX_ARMS;
  const int c = blockIdx.x * KX_RS_CR + lc;
  const bool have = c < d.nCr;
  const KxCreature cr = have ? d.cr[c] : KxCreature{};
  const int A = have ? cr.narms : 0;
  // kraken.js:24-49: segLen, shoulderLimit, jointLimit, ARM_REACH
  const double kPi = 3.141592653589793, seg_len = 13.5;
  const double shoulder_lim = kPi * 0.62, joint_lim = kPi * 0.42, reach = 24.0 + 7.0 * seg_len;
  float *o = out + (size_t)(have ? c : 0) * stride;
  if (A > 0 && a < A) {
    // this arm's inputs (core.js:140-160), each computed once
    const int *so = d.arm_seg_off + cr.arm_off;
    const int m = d.arm_mantle[cr.arm_off];
    const double x = d.px[m], y = d.py[m], angle = d.ang[m];
    const double ca = cos(-angle), sa = sin(-angle);
    const double ds = (double)d.ang[d.arm_segs[so[a]]] - angle;
    const double shoulder = atan2(sin(ds), cos(ds));
    double cs = 0.0;
    int nj = 0;
    for (int j = so[a] + 1; j < so[a + 1]; j++) {
      const double dd = (double)d.ang[d.arm_segs[j]] - (double)d.ang[d.arm_segs[j - 1]];
      cs += atan2(sin(dd), cos(dd));
      nj++;
    }
    const double curl = nj ? cs / nj : 0.0;
    const int tb = d.arm_segs[so[a + 1] - 1];
    const double L = seg_len * 0.5, t = d.ang[tb];
    const double tx = (double)d.px[tb] + L * cos(t), ty = (double)d.py[tb] + L * sin(t);
    const int s0 = d.arm_segs[so[a]];
    const double ext = kx_js_min(1.0, hypot(tx - (double)d.px[s0], ty - (double)d.py[s0]) / kx_js_max(1.0, reach));
    const int g = d.arm_grip[cr.arm_off + a];
    const bool brk = d.broken[g] != 0;
    o[a] = (float)kx_cl(shoulder / shoulder_lim, 1.2);
    o[A + a] = (float)kx_cl(curl / joint_lim, 1.2);
    const double dx = tx - x, dy = ty - y;
    o[2 * A + 2 * a] = (float)kx_cl((dx * ca - dy * sa) / reach, 1.2);
    o[2 * A + 2 * a + 1] = (float)kx_cl((dx * sa + dy * ca) / reach, 1.2);
    // the grip's force as _postGripBreak last measured it: |lambda| / h of
    // the step's last substep (xpbd.js:496-504)
    const double f = hypot((double)d.lamx[g], (double)d.lamy[g]) * (1.0 / d.hd);
    const double fz = (f == f && f != 0.0) ? f : 0.0;
    o[4 * A + a] = brk ? 0.0f : (float)kx_js_min(1.0, fz / kx_js_max(1.0, (double)d.bf[g]));
    o[5 * A + a] = brk ? 0.0f : 1.0f;
    s_ext[lc][a] = ext;
    s_curl[lc][a] = curl;
    s_brk[lc][a] = brk;
  }
  __syncthreads();
  if (A <= 0 || a != 0) return;
  // the creature's sums over its arms in arm order, then its mantle values
  const int m = d.arm_mantle[cr.arm_off];
  const double x = d.px[m], y = d.py[m], vx = d.vx[m], vy = d.vy[m], angle = d.ang[m], av = d.av[m];
  const double speed = hypot(vx, vy), ca = cos(-angle), sa = sin(-angle);
  int anchored = 0;
  double ext_sum = 0.0, curl_sum = 0.0;
  for (int k = 0; k < A; k++) {
    ext_sum += s_ext[lc][k];
    curl_sum += s_curl[lc][k];
    if (!s_brk[lc][k]) anchored++;
  }
  int i = 6 * A;
  const double pvx = d.prev_v[2 * c], pvy = d.prev_v[2 * c + 1];
  const double ax = (vx - pvx) / (d.hd * d.S), ay = (vy - pvy) / (d.hd * d.S);
  o[i++] = (float)kx_js_max(-1.0, kx_js_min(1.0, (ax * ca - ay * sa) / 600));
  o[i++] = (float)kx_js_max(-1.0, kx_js_min(1.0, (ax * sa + ay * ca) / 600));
  const double mean_ext = ext_sum / A, mean_curl = curl_sum / A;
  o[i++] = (float)((double)anchored / A);
  o[i++] = (float)mean_ext;
  o[i++] = (float)kx_js_min(1.0, speed / 300);
  o[i++] = (float)kx_js_max(-1.0, kx_js_min(1.0, av / 18));
  // mantle state and the readout's side values
  const double extra[KX_STATE_EXTRA] = {x, y, angle, vx, vy, av, speed, mean_ext, mean_curl, (double)anchored, ax, ay};
  for (int k = 0; k < KX_STATE_EXTRA; k++) o[i++] = (float)extra[k];
  d.prev_v[2 * c] = (float)vx;
  d.prev_v[2 * c + 1] = (float)vy;
}

int kx_launch_read_state(const KxDev &d, float *out, int stride, hipStream_t st) {
  hipLaunchKernelGGL(kx_k_read_state, dim3((d.nCr + KX_RS_CR - 1) / KX_RS_CR), dim3(64), 0, st, d, out, stride);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}

// respawnCreature (src/sim/train.js:188-229) in place: one thread per
// (listed creature, arm), eight creatures per 64-thread block. Arm a's
// thread re-chains the arm from 24 px out along angle + 2*pi*a/A in 13.5 px
// segments turning 0.28 rad each (float64, then rounded to the state's
// binary32 as the reference's upload rounds its float64 pose), zeroes its
// velocities, sets every motor target of the arm to 0 and releases its grip
// (broken, lambdaX/Y 0). Arm 0's thread also places the mantle and clears
// the readout's previous mantle velocity (readProprio's prevVx/prevVy).
// Compliances, other multipliers and reflex cooldowns stay as they were.
// Creature ids are in range and distinct (checked on the host).
__global__ __launch_bounds__(64) void kx_k_respawn(KxDev d, const int *ids, const double *pose, int n) {
  const int li = blockIdx.x * (64 / KX_MAX_ARMS) + threadIdx.x / KX_MAX_ARMS, a = threadIdx.x % KX_MAX_ARMS;
  if (li >= n) return;
  const int c = ids[li];
  const KxCreature cr = d.cr[c];
  const int A = cr.narms;
  if (a >= A) return;
  const double x = pose[3 * li], y = pose[3 * li + 1], angle = pose[3 * li + 2];
  const double kPi = 3.141592653589793, seg_half = 13.5 * 0.5;
  if (a == 0) {
    const int m = d.arm_mantle[cr.arm_off];
    d.px[m] = (float)x; d.py[m] = (float)y; d.ppx[m] = (float)x; d.ppy[m] = (float)y;
    d.vx[m] = 0.0f; d.vy[m] = 0.0f; d.av[m] = 0.0f;
    d.ang[m] = (float)angle; d.angP[m] = (float)angle;
    d.prev_v[2 * c] = 0.0f;
    d.prev_v[2 * c + 1] = 0.0f;
  }
  const double th0 = angle + ((double)a / A) * kPi * 2;
  double jx = x + cos(th0) * 24, jy = y + sin(th0) * 24, th = th0;
  const int *so = d.arm_seg_off + cr.arm_off;
  for (int j = so[a]; j < so[a + 1]; j++) {
    const int b = d.arm_segs[j];
    const double ux = cos(th), uy = sin(th);
    const double cx = jx + ux * seg_half, cy = jy + uy * seg_half;
    d.px[b] = (float)cx; d.py[b] = (float)cy; d.ppx[b] = (float)cx; d.ppy[b] = (float)cy;
    d.vx[b] = 0.0f; d.vy[b] = 0.0f; d.av[b] = 0.0f;
    d.ang[b] = (float)th; d.angP[b] = (float)th;
    jx = cx + ux * seg_half;
    jy = cy + uy * seg_half;
    th += 0.28;
  }
  for (int k = cr.cons_off; k < cr.cons_off + cr.nc; k++)
    if (d.ctype[k] == KX_ANGLE && d.cslot[k] >= 0 && d.cslot[k] / KX_SLOTS_PER_ARM == a) d.target[k] = 0.0f;
  const int g = d.arm_grip[cr.arm_off + a];
  d.broken[g] = 1;
  d.lamx[g] = 0.0f;
  d.lamy[g] = 0.0f;
}

int kx_launch_respawn(const KxDev &d, const int *ids, const double *pose, int n, hipStream_t st) {
  const int per = 64 / KX_MAX_ARMS;
  hipLaunchKernelGGL(kx_k_respawn, dim3((n + per - 1) / per), dim3(64), 0, st, d, ids, pose, n);
  return hipGetLastError() == hipSuccess ? 0 : -5;
}
