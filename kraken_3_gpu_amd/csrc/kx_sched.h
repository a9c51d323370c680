// Morphology templates for the persistent path (see kx_sched.cpp).
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

// Device view of one template: word offsets (relative to `base`) into the
// template pool. `words` words starting at pool[base] are copied to LDS.
struct KxTmpl {
  int base, words;
  int cr_stride;                  // LDS words per creature region of this template (set by the host)
  int nb, nc, np, narms, nmot, ngrip, R, maxbl;
  int o_crec, o_brec, o_prec;     // 16-byte records (see kx_pack_template)
  // o_sched: groups of four rounds, per lane four packed slots
  //          k | A << 8 | B << 16 (-1 idle);
  // o_rtype: round-type bitmask words (bit r set = angle round)
  int o_sched, o_rtype, o_coll, o_ccnt, o_arm_tip, o_arm_grip, o_mot_k, o_mot_slot, o_grip_k, o_grip_arm;
  int o_arm_gslot;                // per arm: the slot of its grip in the grip list
  int fused;                      // 1: the generic kernels solve its rounds fused (set by the host)
  int single;                     // 1: each pin or grip on one lane (only with fused)
};

// Host form of a template, in creature-local indices (constraints in the
// creature's colour order, points grouped by body).
struct TmplHost {
  int nb = 0;
  std::vector<int> type, a, b;            // b = -1 for world anchors
  std::vector<float> lax, lay, rest, mn, mx;
  std::vector<float> im, ii;
  std::vector<int> poff, pcnt;            // per body
  std::vector<float> plx, ply, pr;
  std::vector<int> arm_tip, arm_grip;     // per arm (local body / local constraint)
  std::vector<int> mot_k, mot_slot;       // motors with their control-law slot
  std::vector<int> grip_k, grip_arm;      // every grip, its arm (-1 if none)
};

int kx_pin_capacity(int G);
// single: a pin or grip takes one lane (G per round) instead of a lane pair
void kx_schedule_rounds(const TmplHost &t, int G, std::vector<int> &sched, std::vector<int> &rtype, int &R,
                        bool single = false);
void kx_schedule_collisions(const TmplHost &t, int G, std::vector<int> &lists, std::vector<int> &counts, int &maxbl);
KxTmpl kx_pack_template(const TmplHost &t, int G, std::vector<uint32_t> &pool, bool single = false);
std::string kx_template_key(const TmplHost &t);
