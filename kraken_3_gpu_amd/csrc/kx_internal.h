// Internal layout shared by the host API (kx_api.cpp) and the kernels.
//
// Device layout (HBM, structure of arrays):
//   bodies       desc order; every creature's bodies are contiguous
//   constraints  "internal order": creature-major, colour order inside a
//                creature (the handle order restricted to one creature)
//   points       grouped by body, stable (gpuworld.js:189-207)
//   map          rect list + a uniform cull grid (exact for contacts)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/kxpbd.h"
#include "kx_sched.h"

#define KX_MAX_COLORS 16
#define KX_MAX_ARMS 8
#define KX_STATE_EXTRA 12   // include/kxpbd.h kx_read_state
#define KX_BODY_REGS 8   // persistent path: bodies per lane held in registers (nb <= KX_BODY_REGS * lanes)
#define KX_SLOTS_PER_ARM 16
#define KX_SLOT_ADH 256
#define KX_GRIP_DRAW_KEY 0xFFFFFFFFu   // step key of the once-per-tip grip limit draw
#define KX_WAVE 64

// Per-creature metadata (internal indices).
struct KxCreature {
  int body_off, nb;       // bodies [body_off, body_off+nb)
  int cons_off, nc;       // internal constraints [cons_off, cons_off+nc)
  int pt_off, np;         // grouped points [pt_off, pt_off+np)
  int arm_off, narms;     // arms [arm_off, arm_off+narms) in the arm tables
  int ncol, tmpl;         // non-empty colour count, morphology bucket
  uint16_t coff[KX_MAX_COLORS + 1];  // colour offsets inside the creature block
};

// Device pointers + solver constants, passed by value to every kernel.
struct KxDev {
  // bodies
  float *px, *py, *ang, *vx, *vy, *av, *ppx, *ppy, *angP, *fx, *fy, *tq;
  const float *im, *ii;
  // constraints (internal order)
  const int *ca, *cb;            // global body indices, cb = -1 for world anchors
  const int *ctype;              // KX_PIN / KX_ANGLE / KX_GRIP
  const int *ccr;                // creature of the constraint
  const int *cslot;              // control-law slot (arm*16+m) or -1
  const int *carm;               // arm index inside the creature (grips), else -1
  const float *lax, *lay, *rest, *cmin, *cmax;
  float *lbx, *lby, *comp, *target, *bf;
  int *broken;
  float *lam, *lamx, *lamy;
  const int *col_list;           // handle order -> internal index (per-colour lists)
  // points (grouped by body)
  const float *plx, *ply, *pr;
  const int *bpoff, *bpcnt;      // per body
  // map
  const float4 *rects;
  int nr;
  const int *cell_off;           // [gnx*gny+1]
  const float4 *cell_rects;      // the rect of every cell entry (cell by cell, ascending rect index)
  const uint32_t *cell_bits;     // occupancy bitmap (1 bit per cull cell), then the near bitmap
  int cell_words, near_words;    // words of each bitmap (multiples of 4)
  float gx0, gy0, inv_cs;
  int gnx, gny;
  float cull_r;                  // every rect within cull_r + 1 px of a cull cell is registered in it
  float inv_ns;                  // near-bitmap cells per px (a power-of-two multiple of inv_cs)
  int nnx, nny;                  // near-bitmap grid, same origin as the cull grid
  // creatures / arms
  const KxCreature *cr;
  const int *bcr;                // creature of every body
  float *org;                    // [2*nCr] plain path: integer origin of each creature
  const int *arm_tip;            // global body
  const int *arm_grip;           // internal constraint
  const int *arm_mantle;         // body the arm's shoulder motor starts at (state readout)
  const int *arm_seg_off;        // [nArms+1] into arm_segs: the arm's segments, shoulder to tip
  const int *arm_segs;
  float *prev_v;                 // [2*nCr] mantle velocity at the previous state readout
  double *cd;                    // [nCr * KX_MAX_ARMS]
  uint32_t *events;              // [nsteps * nCr] or null
  const float *tape;             // [rows][nCr][KX_TAPE_WORDS] controls under KX_LAW_TAPE, or null
  // constants
  float h, invH, invH2, linDamp, angDamp, maxSpeed, maxAngSpeed;
  // the least binary32 q whose correctly rounded square root exceeds
  // maxSpeed (-0 when every q qualifies, NaN when none does): for the q of
  // a velocity, sqrtf(q) > maxSpeed exactly when q >= maxSpeed2
  float maxSpeed2;
  double hd;                     // h in f64 (reflex cooldown arithmetic)
  int S, I, nB, nC, nCr, nArms;
  int ncol;                      // non-empty colours in handle order
  int col_off[KX_MAX_COLORS], col_cnt[KX_MAX_COLORS];
  int use_grid;                  // plain path: 0 = linear SDF scan (WGSL shape)
};

// Device copy of the control law (kx_law).
struct KxLaw {
  uint32_t seed;
  double comp;                   // 1/(stiff_min + tone*(stiff_max-stiff_min))
  double shoulder_limit, joint_limit;
  double grip_min, grip_max, slip_max, cooldown, attach_min, release_below;
  double stiff_min, stiff_max;   // tone -> compliance of tape controls
  uint32_t cbase, flags;         // creature key offset, KX_LAW_* flags
};

// Launchers (kx_plain.hip / kx_persist.hip). All asynchronous on `st`.
// step key = (step_base ? *step_base : 0) + step_key
// tape_row: the control-tape row of this step (KX_LAW_TAPE)
int kx_launch_read_state(const KxDev &d, float *out, int stride, hipStream_t st);
// in-place respawn of creatures ids[0..n) at pose[3*i..3*i+3) = x, y, angle
int kx_launch_respawn(const KxDev &d, const int *ids, const double *pose, int n, hipStream_t st);
int kx_plain_step(const KxDev &d, const KxLaw *law, int step_key, const int *step_base, int with_forces,
                  int ev_row, int tape_row, hipStream_t st, int *launches);
int kx_plain_substeps(const KxDev &d, int n, int with_forces, hipStream_t st, int *launches);
// Persistent-path launch description (see kx_persist.hip).
struct KxPersistCfg {
  const int *wave_cr;       // creature ids of every wavefront slot, -1 = empty slot
  const int *wg_tmpl;       // [nwg] template of each workgroup
  const int *wg_lanes;      // [nwg] lanes per creature of each workgroup
  const int *wg_cr_off;     // [nwg] first wave_cr slot of each workgroup (WPB * 64/lanes slots)
  int nwg, max_cpw;         // workgroups; most creatures per wavefront of any workgroup
  int mixed;                // lane counts differ between workgroups: launch kx_k_persist_mix
  const KxTmpl *tmpl;
  const uint32_t *pool;
  int tmpl_words;           // LDS words reserved for the template
  int bm_words;             // LDS words of the cell occupancy bitmap (after the template)
  int map_words;            // LDS words of the staged map (after the bitmap), 0 = map stays in HBM
  int map_cells, map_ids;   // staged map: the rect of each of the map_ids cell entries (16 B each),
                            // then map_cells + 1 16-bit cell offsets
  int cr_words;             // LDS words of the creature regions of the largest workgroup
  int NB, NC, NP;           // per-creature array sizes in LDS (bodies, constraints, points)
  int coll_par;             // 1: collisions through the per-creature candidate list (phase A/B);
                            // 0: each near body folded by its owning lane
  int reg_rounds;           // KX_REG_ROUNDS: the register-slot kernel (one template of that many
                            // rounds at 16 lanes, at most one wavefront per SIMD); 0: the LDS-slot kernel
  int fixed_rounds;         // KX_REG_ROUNDS (with reg_rounds 0): the same fixed schedule with the
                            // slots read from LDS (kx_k_persist_fix)
};
int kx_persist_launch(int G, int wpb, const KxDev &d, const KxLaw *law, int step0, const int *step_base,
                      int nsteps, int ev_row0, int with_forces, const KxPersistCfg &P, hipStream_t st);
int kx_persist_cr_words(int NB, int NC, int NP, int NG, int NA, bool par);
int kx_persist_occupancy(int G, int wpb, size_t lds, bool par);   // G = 0: the mixed-lane kernel
// register-slot kernel: 16 lanes per creature, KX_REG_ROUNDS rounds per iteration
#define KX_REG_ROUNDS 12
// the most rounds of a template whose rounds the generic kernels solve fused
#define KX_FUSED_MAX_ROUNDS 16
// its round types, bit r set = round r solves angle motors: the schedule of the
// 6-arm x 7-segment creature (pins and motors alternate along the mantle chain)
#define KX_REG_RTYPE 0xAAA
// fused rounds f (pin round 2f and motor round 2f + 1) of that schedule whose pins
// all join two bodies (the grips fall in rounds 4, 6 and 10)
#define KX_REG_PINB 0x13
#define KX_REG_BODY_REGS 3   // bodies per lane held in registers by the register-slot kernel (43 of the 6 x 7 creature)
int kx_persist_occupancy_reg(int wpb, size_t lds, bool par);
int kx_bump_counter(int *counter, int delta, hipStream_t st);

// One handle's share of a batched persistent launch: its device state and
// launch tables, its control law (KxLaw) and step range. Workgroup b of the
// batched grid runs local workgroup wgmap[b].y of entry wgmap[b].x.
struct KxBatchEnt {
  KxDev d;
  KxLaw L;
  KxPersistCfg P;
  int has_law, step0, nsteps, pad;
};
// LDS bytes a handle's persistent workgroups need
inline size_t kx_persist_lds_bytes(const KxPersistCfg &P) {
  return ((size_t)P.tmpl_words + P.bm_words + P.map_words + (size_t)P.cr_words) * 4;
}
// every entry of one batched launch has the same collision fold `par`
int kx_persist_launch_batch(int wpb, int par, const KxBatchEnt *ents_dev, const int2 *wgmap_dev, int nwg, size_t lds,
                            hipStream_t st);

// What a batched launch needs of a handle (kx_api.cpp); fails unless the
// handle runs the persistent path.
struct kx_handle;
struct KxHandleView {
  KxDev d;
  KxPersistCfg P;
  int wpb;
  hipStream_t st;
  size_t bytes;             // device memory the handle's state arrays hold
};
int kx_handle_view(kx_handle *h, KxHandleView *v);
// Fixes a handle's collision fold (0 = by map density, 1 = parallel over the
// creature's lanes, 2 = by the owning lane) and rebuilds its persistent
// configuration.
int kx_set_coll_mode(kx_handle *h, int mode);
// sets the thread's kx_last_error message and returns code
int kx_error(int code, const char *msg);
// kx_rollout with the event words in `ev_dev` (nsteps x creatures words,
// zeroed by the caller) when events_host is given; allocates nothing
int kx_rollout_into(kx_handle *h, const kx_law *law, int32_t step0, int32_t nsteps, uint32_t *ev_dev,
                    uint32_t *events_host);
// the device a handle was created on, its creature count and stream
int kx_handle_basics(const kx_handle *h, int *device, int *n_creatures, hipStream_t *st);
// Device side of one policy step (kx_api.cpp): reports the creature count and
// the arm count every creature shares; with `launch`, also launches the state
// readout into the handle's device readout buffer (stride floats per
// creature), makes sure the control tape holds at least one row, and returns
// both buffers and the handle's stream. Returns the floats per creature the
// readout writes, or a negative code (nothing launched).
int kx_policy_io(kx_handle *h, int32_t stride, float **state_dev, float **tape_dev, hipStream_t *st, int32_t *n_cr,
                 int32_t *arms, int launch);
