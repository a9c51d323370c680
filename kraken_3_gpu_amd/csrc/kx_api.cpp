// Host side of libkxpbd: the C ABI declared in include/kxpbd.h.
//
// kx_create mirrors gpuworld.create (src/gpu/gpuworld.js:34-171): sanity
// guards (writeStatic 224-229, upload 269-271), colouring (ensureColored,
// src/physics/xpbd.js:108-126), point grouping (groupCollPoints 189-207).
// On top of that it derives the per-creature layout the persistent kernel
// needs (connected components of the constraint graph) and the exact cull
// grid for the map.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/kxpbd.h"
#include "kx_internal.h"
#include "kx_sched.h"

#include <map>

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

extern "C" const char *kx_last_error(void) { return g_err.c_str(); }

int kx_error(int code, const char *msg) {
  g_err = msg;
  return code;
}
extern "C" int kx_abi_version(void) { return KXPBD_ABI_VERSION; }

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) return fail(KX_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define TRY(x)            \
  do {                    \
    int r_ = (x);         \
    if (r_) return r_;    \
  } while (0)

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

struct kx_handle {
  int nB = 0, nC = 0, nP = 0, nR = 0, nCr = 0, nArms = 0;
  kx_params prm{};
  int path = KX_PATH_PERSISTENT, G = 16, wpb = 1;
  int n_cell_ids = 0;
  bool persist_ok = false, has_creatures = false, forces_pending = false;
  std::string persist_why;
  int maxb = 0, maxc = 0, maxp = 0;
  // host tables
  std::vector<int> ord;        // handle order -> desc index
  std::vector<int> colour;     // handle order colour
  std::vector<int> h2i;        // handle order -> internal index
  uint32_t colorOff[KX_MAX_COLORS] = {}, colorCnt[KX_MAX_COLORS] = {};
  std::vector<KxCreature> cr;
  std::vector<int> ctype_i;    // internal order types (host copy)
  // device
  KxDev d{};
  std::vector<DevBuf> bufs;
  int *step_base = nullptr;
  uint32_t *ev_dev = nullptr;
  size_t ev_cap = 0;
  float *tape_dev = nullptr;     // control tape (KX_LAW_TAPE), tape_rows rows valid
  size_t tape_cap = 0;
  int tape_rows = 0;
  hipStream_t st = nullptr;
  int dev = 0;                   // the device current at kx_create
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int last_launches = 0;
  bool timed = false;
  // graph cache
  hipGraphExec_t gexec = nullptr;
  int g_nsteps = -1, g_path = -1;
  KxLaw g_law{};
  // persistent path: morphology templates and the wave map for shape (G, wpb)
  std::vector<TmplHost> tmpl_host;
  std::vector<int> cr_tmpl;
  std::vector<int> tmpl_rounds;
  KxPersistCfg pc{};
  void *pc_bufs[4] = {nullptr, nullptr, nullptr, nullptr};
  int pinfo[13] = {};    // see kx_persist_info
  std::vector<int> tmpl_lanes;   // per template lanes per creature, 0 = the handle's G
  int launch_G = 16;             // lane count the uniform kernel is launched with
  int coll_mode = 0;             // collision fold: 0 = by map density, 1 = parallel, 2 = owning lane
  int reg_mode = 1;              // 1: register-held round slots where the launch allows them
  std::string state_why;         // why the state readout is unavailable (empty: available)
  float *state_dev = nullptr;    // readout buffer [nCr * state_stride], state_cap floats allocated
  int state_stride = 0;
  size_t state_cap = 0;
  void *resp_dev = nullptr;      // kx_respawn staging: ids then poses, resp_cap bytes
  size_t resp_cap = 0;
  double near_frac = 0.0;        // fraction of the near bitmap set (build_grid)
  // scratch host staging
  std::vector<float> hf;
  std::vector<int> hi;
};

template <class T>
static int dalloc(kx_handle *h, T **p, size_t n) {
  size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
  void *q = nullptr;
  HIPCHK(hipMalloc(&q, bytes));
  HIPCHK(hipMemset(q, 0, bytes));
  h->bufs.push_back({q, bytes});
  *p = (T *)q;
  return 0;
}
// A handle buffer that grows in place of being duplicated: when `need`
// bytes exceed `*cap`, the old allocation is released (after the handle's
// queued work) and dropped from the handle's accounting, and a larger one
// takes its place. Contents are not kept.
static int regrow(kx_handle *h, void **p, size_t *cap, size_t need) {
  if (*p && need <= *cap) return 0;
  if (*p) {
    HIPCHK(hipStreamSynchronize(h->st));
    for (size_t i = 0; i < h->bufs.size(); i++)
      if (h->bufs[i].p == *p) { h->bufs.erase(h->bufs.begin() + (long)i); break; }
    HIPCHK(hipFree(*p));
    *p = nullptr;
    *cap = 0;
  }
  char *q = nullptr;
  TRY(dalloc(h, &q, need));
  *p = q;
  *cap = need;
  return 0;
}

template <class T>
static int dput(T *dst, const T *src, size_t n) {
  if (n) HIPCHK(hipMemcpy(dst, src, n * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

// ── colouring (xpbd.js:98-126) ──────────────────────────────────────
static int colour_constraints(int nB, int nC, const int32_t *ca, const int32_t *cb, const int32_t *given,
                              std::vector<int> &col) {
  col.assign(nC, -1);
  std::vector<uint32_t> mask(std::max(1024, nB * 2), 0u);
  for (int k = 0; k < nC; k++) col[k] = given ? given[k] : -1;
  // pre-existing colours occupy their bodies first, like assignColor on an
  // already-coloured world (only uncoloured constraints are assigned)
  for (int k = 0; k < nC; k++) {
    if (col[k] < 0) {
      uint32_t used = mask[ca[k]] | (cb[k] >= 0 ? mask[cb[k]] : 0u);
      int c = 0;
      while (c < 32 && (used & (1u << c))) c++;
      if (c >= KX_MAX_COLORS) return fail(KX_EINVAL, "constraint %d needs colour %d >= %d", k, c, KX_MAX_COLORS);
      mask[ca[k]] |= 1u << c;
      if (cb[k] >= 0) mask[cb[k]] |= 1u << c;
      col[k] = c;
    } else if (col[k] >= KX_MAX_COLORS) {
      return fail(KX_EINVAL, "constraint %d has colour %d >= %d", k, col[k], KX_MAX_COLORS);
    }
  }
  return 0;
}

extern "C" int kx_plan_rounds(int32_t nb, int32_t nc, const int32_t *ca, const int32_t *cb, const int32_t *ct,
                              int32_t G, int32_t *sched_out, int32_t *rtype_out, int32_t max_rounds) {
  const bool single = (G & KX_PLAN_SINGLE_LANE) != 0;
  G &= ~KX_PLAN_SINGLE_LANE;
  if (nb < 1 || nc < 0 || G < 1 || G > 64) return fail(KX_EINVAL, "bad arguments");
  TmplHost t;
  t.nb = nb;
  for (int k = 0; k < nc; k++) {
    if (ca[k] < 0 || ca[k] >= nb || cb[k] < -1 || cb[k] >= nb) return fail(KX_EINVAL, "constraint %d out of range", k);
    t.type.push_back(ct[k]); t.a.push_back(ca[k]); t.b.push_back(cb[k]);
  }
  std::vector<int> sched, rtype;
  int R = 0;
  kx_schedule_rounds(t, G, sched, rtype, R, single);
  if (R > max_rounds) return fail(KX_EINVAL, "needs %d rounds > %d", R, max_rounds);
  for (size_t i = 0; i < sched.size(); i++) sched_out[i] = sched[i];
  for (int r = 0; r < R; r++) rtype_out[r] = rtype[r];
  return R;
}

extern "C" int kx_color_constraints(int32_t nB, int32_t nC, const int32_t *ca, const int32_t *cb,
                                    const int32_t *color_in, int32_t *color_out, int32_t *order_out) {
  if (nB < 1 || nC < 0 || (nC && (!ca || !cb))) return fail(KX_EINVAL, "bad arguments");
  for (int k = 0; k < nC; k++)
    if (ca[k] < 0 || ca[k] >= nB || cb[k] < -1 || cb[k] >= nB)
      return fail(KX_EINVAL, "constraint %d: body index %d/%d outside %d", k, ca[k], cb[k], nB);
  std::vector<int> col;
  TRY(colour_constraints(nB, nC, ca, cb, color_in, col));
  std::vector<int> ord(nC);
  std::iota(ord.begin(), ord.end(), 0);
  std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return col[x] < col[y]; });
  for (int k = 0; k < nC; k++) {
    if (color_out) color_out[k] = col[k];
    if (order_out) order_out[k] = ord[k];
  }
  return 0;
}

extern "C" int kx_group_points(int32_t nP, const int32_t *p_body, int32_t nB, int32_t *order, int32_t *range) {
  std::vector<int32_t> cnt(nB, 0);
  for (int i = 0; i < nP; i++) {
    int b = p_body[i];
    if (b < 0 || b >= nB) return fail(KX_EINVAL, "collPt %d refers to body %d of %d", i, b, nB);
    cnt[b]++;
  }
  std::vector<int32_t> off(nB, 0);
  int acc = 0;
  for (int b = 0; b < nB; b++) { off[b] = acc; acc += cnt[b]; }
  if (acc != nP) return fail(KX_EINVAL, "collRange covers %d of %d points", acc, nP);
  std::vector<int32_t> cur(off);
  for (int i = 0; i < nP; i++) order[cur[p_body[i]]++] = i;
  for (int b = 0; b < nB; b++) { range[b] = off[b]; range[nB + b] = cnt[b]; }
  return 0;
}

// union-find
static int uf_find(std::vector<int> &p, int x) {
  while (p[x] != x) { p[x] = p[p[x]]; x = p[x]; }
  return x;
}

// The least binary32 q with sqrtf(q) > m (sqrtf correctly rounded, so
// monotonic in q): the speed clamp's test sqrtf(vx*vx + vy*vy) > maxSpeed is
// then q >= threshold, and the square root is taken only to clamp. -0 when
// every q >= 0 qualifies (m < 0), NaN when no q does (m NaN or +inf).
static float speed_threshold(float m) {
  if (std::isnan(m) || m == INFINITY) return NAN;
  if (m < 0.0f) return -0.0f;
  float q = m * m;   // within a few ulps of the threshold (inf for a huge m)
  while (q > 0.0f && std::sqrt(std::nextafter(q, 0.0f)) > m) q = std::nextafter(q, 0.0f);
  while (!(std::sqrt(q) > m)) q = std::nextafter(q, INFINITY);
  return q;
}

// ── exact cull grid over the rect list ──────────────────────────────
// Also two cell bitmaps, each padded to a multiple of four words and stored
// back to back: occupancy (cull cells with a registered rect) and the near
// bitmap, on a grid with the cull grid's origin and cells of cs / m (m = 2
// when that costs at most 256 words, else 1). A near cell's bit is set when
// the distance from the cell to some rect's box is below reach + rmax plus a
// margin, where reach bounds every collision point's distance from its body
// centre and rmax every point's radius. A point p of a body centred at c in a
// clear cell is then farther than rmax from every rect, so no query of p
// corrects it (a correction needs the fold's distance below the point's
// radius, and the fold's nearest point lies in the rect's box whatever the
// signs of its width and height). The margin (0.05 px + fs / 1000) covers the
// binary32 rounding of the device's cell expression and of the world point; a
// centre outside the grid clamps to the edge cell, whose distance to every
// rect (all inside the grid) is no larger than the centre's.
static int build_grid(kx_handle *h, const float *rects, int nR, float rmax, float reach) {
  float half = 0.0f;
  float x0 = INFINITY, y0 = INFINITY, x1 = -INFINITY, y1 = -INFINITY;
  for (int i = 0; i < nR; i++) {
    const float *r = rects + 4 * i;
    half = std::max(half, 0.5f * std::min(std::fabs(r[2]), std::fabs(r[3])));
    x0 = std::min(x0, r[0]); y0 = std::min(y0, r[1]);
    x1 = std::max(x1, r[0] + r[2]); y1 = std::max(y1, r[1] + r[3]);
  }
  float R = std::max(half, rmax) + 2.0f;
  float cs = std::max(32.0f, 2.0f * R);
  int gnx = 1, gny = 1;
  std::vector<int> off(2, 0), ids;
  float gx0 = 0, gy0 = 0;
  if (nR > 0) {
    gx0 = std::floor(x0 - R - cs);
    gy0 = std::floor(y0 - R - cs);
    gnx = (int)std::ceil((x1 + R + cs - gx0) / cs) + 1;
    gny = (int)std::ceil((y1 + R + cs - gy0) / cs) + 1;
    if ((long)gnx * gny > (1L << 24)) return fail(KX_EINVAL, "map extent too large for the cull grid");
    std::vector<std::vector<int>> cells((size_t)gnx * gny);
    for (int i = 0; i < nR; i++) {
      const float *r = rects + 4 * i;
      float ax0 = r[0] - R - 1.0f, ax1 = r[0] + r[2] + R + 1.0f;
      float ay0 = r[1] - R - 1.0f, ay1 = r[1] + r[3] + R + 1.0f;
      int ix0 = std::max(0, (int)std::floor((ax0 - gx0) / cs)), ix1 = std::min(gnx - 1, (int)std::floor((ax1 - gx0) / cs));
      int iy0 = std::max(0, (int)std::floor((ay0 - gy0) / cs)), iy1 = std::min(gny - 1, (int)std::floor((ay1 - gy0) / cs));
      for (int iy = iy0; iy <= iy1; iy++)
        for (int ix = ix0; ix <= ix1; ix++) cells[(size_t)iy * gnx + ix].push_back(i);
    }
    off.assign((size_t)gnx * gny + 1, 0);
    for (size_t c = 0; c < cells.size(); c++) {
      off[c + 1] = off[c] + (int)cells[c].size();
      ids.insert(ids.end(), cells[c].begin(), cells[c].end());
    }
  }
  auto words4 = [](size_t cells) { return ((cells + 31) / 32 + 3) / 4 * 4; };
  const size_t W4 = words4((size_t)gnx * gny);
  const int m = words4((size_t)4 * gnx * gny) <= 256 ? 2 : 1;
  const int nnx = m * gnx, nny = m * gny;
  const size_t N4 = words4((size_t)nnx * nny);
  std::vector<uint32_t> bits(W4 + N4, 0u);
  for (size_t c = 0; c + 1 < off.size(); c++)
    if (off[c + 1] > off[c]) bits[c >> 5] |= 1u << (c & 31);
  {
    const double fs = (double)cs / m;
    const double D = (double)reach + (double)rmax + 0.05 + 1e-3 * fs;
    uint32_t *nb = bits.data() + W4;
    for (int i = 0; i < nR; i++) {
      const float *r = rects + 4 * i;
      const double bx0 = std::min((double)r[0], (double)r[0] + r[2]), bx1 = std::max((double)r[0], (double)r[0] + r[2]);
      const double by0 = std::min((double)r[1], (double)r[1] + r[3]), by1 = std::max((double)r[1], (double)r[1] + r[3]);
      if (!(std::isfinite(bx0) && std::isfinite(bx1) && std::isfinite(by0) && std::isfinite(by1))) {
        for (size_t k = 0; k < N4; k++) nb[k] = ~0u;      // unknown geometry: every body is a candidate
        break;
      }
      const int ix0 = std::max(0, (int)std::floor((bx0 - D - gx0) / fs) - 1);
      const int ix1 = std::min(nnx - 1, (int)std::floor((bx1 + D - gx0) / fs) + 1);
      const int iy0 = std::max(0, (int)std::floor((by0 - D - gy0) / fs) - 1);
      const int iy1 = std::min(nny - 1, (int)std::floor((by1 + D - gy0) / fs) + 1);
      for (int iy = iy0; iy <= iy1; iy++)
        for (int ix = ix0; ix <= ix1; ix++) {
          const double cx0 = gx0 + ix * fs, cy0 = gy0 + iy * fs;
          const double dx = std::max(0.0, std::max(bx0 - (cx0 + fs), cx0 - bx1));
          const double dy = std::max(0.0, std::max(by0 - (cy0 + fs), cy0 - by1));
          if (dx * dx + dy * dy < D * D) {
            const size_t e = (size_t)iy * nnx + ix;
            nb[e >> 5] |= 1u << (e & 31);
          }
        }
    }
  }
  uint32_t *dbits;
  TRY(dalloc(h, &dbits, bits.size()));
  TRY(dput(dbits, bits.data(), bits.size()));
  h->d.cell_bits = dbits;
  h->d.cell_words = (int)W4;
  h->d.near_words = (int)N4;
  {
    size_t set = 0;
    for (size_t k = 0; k < N4; k++) set += (size_t)__builtin_popcount(bits[W4 + k]);
    h->near_frac = nnx * nny > 0 ? (double)set / ((double)nnx * nny) : 0.0;
  }
  h->d.inv_ns = (float)m / cs;
  h->d.nnx = nnx;
  h->d.nny = nny;
  // each cell entry holds its rect itself, so a query reads a cell's
  // offsets and then its rects (two dependent loads, not three)
  std::vector<float4> crects(ids.size());
  for (size_t k = 0; k < ids.size(); k++) {
    const float *r = rects + 4 * (size_t)ids[k];
    crects[k] = make_float4(r[0], r[1], r[2], r[3]);
  }
  int *doff;
  float4 *dcr;
  TRY(dalloc(h, &doff, off.size()));
  TRY(dalloc(h, &dcr, crects.size()));
  TRY(dput(doff, off.data(), off.size()));
  if (!crects.empty()) TRY(dput((float *)dcr, (const float *)crects.data(), 4 * crects.size()));
  h->d.cell_off = doff; h->d.cell_rects = dcr;
  h->n_cell_ids = (int)ids.size();
  h->d.gx0 = gx0; h->d.gy0 = gy0; h->d.inv_cs = 1.0f / cs; h->d.gnx = gnx; h->d.gny = gny;
  h->d.cull_r = R;
  return 0;
}

// Lanes per creature of template t: its own setting, else the handle's.
static int tmpl_lanes_of(const kx_handle *h, size_t t) {
  return t < h->tmpl_lanes.size() && h->tmpl_lanes[t] ? h->tmpl_lanes[t] : h->G;
}

// Whether a packed template's rounds can run fused: an even number of at
// most KX_FUSED_MAX_ROUNDS rounds alternating pins (even rounds) and motors
// (odd rounds), every motor between two bodies with a finite rest angle and
// limits that are not inverted (no min > max: the fused motor solve takes the
// limit excess as a sum of two one-sided terms), each on the even lane of the
// pin before it (its lane when pins take single lanes) with the same two
// bodies (kx_pack_template places it there when it can).
static bool fused_schedule(const KxTmpl &k, const TmplHost &th, const std::vector<uint32_t> &pool, int G) {
  if (k.R < 2 || (k.R & 1) || k.R > KX_FUSED_MAX_ROUNDS) return false;
  if (pool[(size_t)k.base + k.o_rtype] != (0xAAAAAAAAu & ((1u << k.R) - 1u))) return false;
  for (size_t j = 0; j < th.type.size(); j++)
    if (th.type[j] == KX_ANGLE && (th.b[j] < 0 || !std::isfinite(th.rest[j]) || th.mn[j] > th.mx[j])) return false;
  const uint32_t *desc = pool.data() + k.base + k.o_sched;
  auto slot = [&](int r, int lane) { return (int)desc[((r >> 2) * G + lane) * 4 + (r & 3)]; };
  for (int r = 1; r < k.R; r += 2)
    for (int l = 0; l < G; l++) {
      const int a = slot(r, l);
      if (a < 0) continue;
      const int p = slot(r - 1, l);
      if (((l & 1) && !k.single) || p < 0 || (p >> 8) != (a >> 8)) return false;
    }
  return true;
}

// Pack every template's schedule for its lane count G_t and map creatures
// onto workgroups: WPB wavefronts of 64/G_t creatures of one template each.
static int build_persist(kx_handle *h) {
  const int WPB = h->wpb;
  std::vector<uint32_t> pool;
  std::vector<KxTmpl> T;
  int words = 0, NB = 1, NC = 1, NP = 1, NG = 0, NA = 0;
  h->tmpl_rounds.clear();
  std::vector<std::vector<int>> members(h->tmpl_host.size());
  for (int c = 0; c < h->nCr; c++) members[h->cr_tmpl[c]].push_back(c);
  int dev = 0, ncu = 1;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  // Collision fold. On an open map (under half of the near bitmap set) a
  // body is near a rect in few iterations and a correcting point is rare, so
  // the near bodies' points are tested in parallel over each creature's
  // lanes and only a body with a correcting point is re-run in order (the
  // parallel fold: a candidate list in LDS); its worst case per wave stays
  // short, which is what a launch of one wave per SIMD waits for. On a dense
  // map (corridors) most bodies are near in most iterations and many
  // correct, so each near body is folded once, in order, by its owning lane
  // (the owner fold: no list, no re-run). A handle may fix the fold
  // (coll_mode 1 = parallel, 2 = owner); both are exact.
  const bool par = h->coll_mode == 1 || (h->coll_mode == 0 && h->near_frac < 0.5);
  for (size_t t = 0; t < h->tmpl_host.size(); t++) {
    const int Gt = tmpl_lanes_of(h, t);
    // each lane holds the start poses and velocities of its bodies in registers
    if (h->tmpl_host[t].nb > KX_BODY_REGS * Gt)
      return fail(KX_ENOTSUP, "a creature of %d bodies needs at least %d lanes (at most %d bodies per lane)",
                  h->tmpl_host[t].nb, (h->tmpl_host[t].nb + KX_BODY_REGS - 1) / KX_BODY_REGS, KX_BODY_REGS);
    // The template's rounds, fused when the schedule allows: pins on lane
    // pairs first (16 or 32 lanes; the kernels specialised to the default
    // creature take only that form), else on single lanes, which hold twice
    // the pins per round (an 8-lane schedule then alternates too).
    const TmplHost &th = h->tmpl_host[t];
    std::vector<uint32_t> scr;
    KxTmpl k = kx_pack_template(th, Gt, scr, false);
    k.fused = 0;
    if (h->reg_mode && (Gt == 16 || Gt == 32) && fused_schedule(k, th, scr, Gt)) {
      k.fused = 1;
    } else if (h->reg_mode && (Gt == 8 || Gt == 16 || Gt == 32)) {
      std::vector<uint32_t> scr1;
      KxTmpl k1 = kx_pack_template(th, Gt, scr1, true);
      if (fused_schedule(k1, th, scr1, Gt)) { k = k1; k.fused = 1; scr.swap(scr1); }
    }
    while (pool.size() % 4) pool.push_back(0u);
    k.base = (int)pool.size();
    pool.insert(pool.end(), scr.begin(), scr.end());
    // creature regions: 16-byte records, regions offset by 16 bytes modulo
    // 128 so neighbouring creatures' same records fall in different banks
    k.cr_stride = (kx_persist_cr_words(k.nb, k.nc, k.np, k.ngrip, k.narms, par) + 31) / 32 * 32 + 4;
    T.push_back(k);
    words = std::max(words, k.words);
    NB = std::max(NB, k.nb); NC = std::max(NC, k.nc); NP = std::max(NP, k.np); NG = std::max(NG, k.ngrip);
    NA = std::max(NA, k.narms);
    h->tmpl_rounds.push_back(k.R);
  }
  std::vector<int> wave_cr, wg(0);
  std::vector<int> wg_tmpl, wg_lanes, wg_off;
  int max_cpw = 1, lanes0 = -1, cr_words = 0;
  bool mixed = false;
  // Workgroups are issued longest first (per-wave cost ~ rounds per
  // iteration plus collision bodies per lane), so the dispatcher's dynamic
  // assignment ends on short workgroups rather than on the largest
  // morphologies; consecutive workgroups of similar cost also spread evenly
  // over the XCDs, which take blockIdx round-robin.
  std::vector<int> order(T.size());
  for (size_t t = 0; t < T.size(); t++) order[t] = (int)t;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return 8 * T[x].R + 4 * T[x].maxbl > 8 * T[y].R + 4 * T[y].maxbl;
  });
  for (int t : order) {
    auto &m = members[t];
    if (m.empty()) continue;
    const int Gt = tmpl_lanes_of(h, t), CPW = KX_WAVE / Gt;
    if (lanes0 < 0) lanes0 = Gt;
    mixed = mixed || Gt != lanes0;
    max_cpw = std::max(max_cpw, CPW);
    const int per_wg = WPB * CPW;
    cr_words = std::max(cr_words, per_wg * T[t].cr_stride);
    const int nwg_t = (int)((m.size() + per_wg - 1) / per_wg);
    for (int b = 0; b < nwg_t; b++) {
      wg_tmpl.push_back((int)t);
      wg_lanes.push_back(Gt);
      wg_off.push_back((int)wave_cr.size());
      for (int j = 0; j < per_wg; j++) {
        size_t idx = (size_t)b * per_wg + j;
        wave_cr.push_back(idx < m.size() ? m[idx] : -1);
      }
    }
  }
  const int nwg = (int)wg_tmpl.size();
  const int Gk = mixed ? 0 : (lanes0 > 0 ? lanes0 : h->G);    // kernel: 0 = mixed-lane kernel
  int bmw = h->d.cell_words + h->d.near_words;   // occupancy + near bitmaps
  if ((size_t)bmw * 4 > 32 * 1024) bmw = 0;      // huge grids: no LDS bitmaps
  // Stage the whole map (rects of the cell entries, 16-bit cell offsets) in
  // LDS when that costs no occupancy the launch can use: with nwg workgroups
  // on nCU CUs, at most ceil(nwg / nCU) of them are ever resident on one CU.
  int mcells = h->d.gnx * h->d.gny, mids = h->n_cell_ids;
  int mapw = 4 * mids + ((mcells + 2) / 2 + 3) / 4 * 4;
  if (mids > 65535) mapw = 0;                 // offsets are staged as 16-bit
  const size_t base = ((size_t)words + bmw + (size_t)cr_words) * 4;
  if (base > 160 * 1024)
    return fail(KX_ENOTSUP, "persistent shape (%d creatures per wave) x %d waves needs %zu B of LDS (> 160 KiB)",
                max_cpw, WPB, base);
  const int need = (std::max(nwg, 1) + std::max(ncu, 1) - 1) / std::max(ncu, 1);
  const int occ0 = kx_persist_occupancy(Gk, WPB, base, par);
  const size_t with_map = base + (size_t)mapw * 4;
  const int occ1 = with_map <= 160 * 1024 ? kx_persist_occupancy(Gk, WPB, with_map, par) : 0;
  if (!mapw || std::min(occ1, need) < std::min(occ0, need)) mapw = 0;
  const int stride = (kx_persist_cr_words(NB, NC, NP, NG, NA, par) + 31) / 32 * 32 + 4;   // largest region (reported)
  const size_t lds = base + (size_t)mapw * 4;
  h->pinfo[0] = (int)lds; h->pinfo[1] = nwg; h->pinfo[2] = mapw ? occ1 : occ0; h->pinfo[3] = need;
  h->pinfo[4] = mapw; h->pinfo[5] = bmw; h->pinfo[6] = words; h->pinfo[7] = stride;
  h->pinfo[8] = mixed ? 1 : 0;
  h->pinfo[9] = par ? KX_FOLD_PARALLEL : KX_FOLD_OWNER;
  // The fixed schedule: every template fused (fused_schedule) with
  // KX_REG_ROUNDS rounds at 16 lanes, the pins of the fused rounds in
  // KX_REG_PINB all with a second body, at most KX_REG_BODY_REGS bodies per
  // lane. With it, register-held round slots when every workgroup is resident
  // at once at one wavefront per SIMD (the slots take the register file, so a
  // SIMD holds no second wave), else the slots read from LDS. Without it, the
  // generic kernels solve each fused template's rounds fused (KxTmpl::fused:
  // pins on lane pairs at 16 or 32 lanes, or on single lanes at 8, 16 or 32)
  // and every other template's rounds one at a time.
  int reg = 0, fixed = 0;
  if (h->reg_mode && !mixed && Gk == 16) {
    bool all = true;
    for (size_t t = 0; t < T.size(); t++) {
      if (members[t].empty()) continue;
      if (!T[t].fused || T[t].single || T[t].R != KX_REG_ROUNDS || T[t].nb > KX_REG_BODY_REGS * 16) all = false;
      if (all) {
        const uint32_t *desc = pool.data() + T[t].base + T[t].o_sched;
        auto slot = [&](int r, int lane) { return (int)desc[((r >> 2) * 16 + lane) * 4 + (r & 3)]; };
        for (int f = 0; f < KX_REG_ROUNDS / 2; f++)
          if ((KX_REG_PINB >> f) & 1)
            for (int l = 0; l < 16; l++) {
              const int p = slot(2 * f, l);
              if (p >= 0 && ((p >> 16) & 255) == 255) all = false;
            }
      }
    }
    const size_t lds_used = base + (size_t)mapw * 4;
    if (all && h->reg_mode == 1 && (long)nwg * WPB <= 4L * ncu && kx_persist_occupancy_reg(WPB, lds_used, par) >= need)
      reg = KX_REG_ROUNDS;
    else if (all)
      fixed = KX_REG_ROUNDS;
  }
  h->pinfo[10] = reg;
  h->pinfo[11] = fixed;
  int fused_cr = 0;
  if (!reg && !fixed)
    for (size_t t = 0; t < T.size(); t++) fused_cr += T[t].fused ? (int)members[t].size() : 0;
  h->pinfo[12] = fused_cr;
  std::vector<int> tables;
  tables.insert(tables.end(), wg_tmpl.begin(), wg_tmpl.end());
  tables.insert(tables.end(), wg_lanes.begin(), wg_lanes.end());
  tables.insert(tables.end(), wg_off.begin(), wg_off.end());
  for (void *&p : h->pc_bufs) { if (p) (void)hipFree(p); p = nullptr; }
  HIPCHK(hipMalloc(&h->pc_bufs[0], std::max<size_t>(pool.size(), 1) * 4));
  HIPCHK(hipMalloc(&h->pc_bufs[1], std::max<size_t>(T.size(), 1) * sizeof(KxTmpl)));
  HIPCHK(hipMalloc(&h->pc_bufs[2], std::max<size_t>(wave_cr.size(), 1) * 4));
  HIPCHK(hipMalloc(&h->pc_bufs[3], std::max<size_t>(tables.size(), 1) * 4));
  HIPCHK(hipMemcpy(h->pc_bufs[0], pool.data(), pool.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->pc_bufs[1], T.data(), T.size() * sizeof(KxTmpl), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->pc_bufs[2], wave_cr.data(), wave_cr.size() * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(h->pc_bufs[3], tables.data(), tables.size() * 4, hipMemcpyHostToDevice));
  KxPersistCfg &P = h->pc;
  P.pool = (const uint32_t *)h->pc_bufs[0];
  P.tmpl = (const KxTmpl *)h->pc_bufs[1];
  P.wave_cr = (const int *)h->pc_bufs[2];
  P.wg_tmpl = (const int *)h->pc_bufs[3];
  P.wg_lanes = P.wg_tmpl + nwg;
  P.wg_cr_off = P.wg_tmpl + 2 * nwg;
  P.nwg = nwg;
  P.max_cpw = max_cpw;
  P.mixed = mixed ? 1 : 0;
  P.tmpl_words = words;
  P.bm_words = bmw;
  P.map_words = mapw;
  P.map_cells = mcells;
  P.map_ids = mids;
  P.cr_words = cr_words;
  P.NB = NB; P.NC = NC; P.NP = NP;
  P.coll_par = par ? 1 : 0;
  P.reg_rounds = reg;
  P.fixed_rounds = fixed;
  h->launch_G = mixed ? h->G : (lanes0 > 0 ? lanes0 : h->G);
  return 0;
}

static void destroy_impl(kx_handle *h) {
  if (!h) return;
  if (h->st) (void)hipStreamSynchronize(h->st);
  if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
  for (auto &b : h->bufs) (void)hipFree(b.p);
  for (void *p : h->pc_bufs) if (p) (void)hipFree(p);
  if (h->ev_dev) (void)hipFree(h->ev_dev);
  if (h->tape_dev) (void)hipFree(h->tape_dev);
  if (h->e0) (void)hipEventDestroy(h->e0);
  if (h->e1) (void)hipEventDestroy(h->e1);
  if (h->st) (void)hipStreamDestroy(h->st);
  delete h;
}

extern "C" void kx_destroy(kx_handle *h) { destroy_impl(h); }

static int create_impl(const kx_world_desc *w, const kx_creature_desc *cd, const kx_params *p, kx_handle *h) {
  const int nB = w->n_bodies, nC = w->n_constraints, nP = w->n_points, nR = w->n_rects;
  // sanity guards (gpuworld.js:224-229)
  if (!(p->iterations >= 1 && p->iterations <= 64)) return fail(KX_EINVAL, "iterations implausible: %d", p->iterations);
  if (!(p->substeps >= 1 && p->substeps <= 64)) return fail(KX_EINVAL, "substeps implausible: %d", p->substeps);
  if (!(nR >= 1 && nR <= 100000)) return fail(KX_EINVAL, "nRects implausible: %d", nR);
  if (!(nB >= 1 && nB <= 4000000)) return fail(KX_EINVAL, "nBodies implausible: %d", nB);
  if (!(p->h > 0 && p->h < 1)) return fail(KX_EINVAL, "h implausible: %g", p->h);
  if (nC < 0 || nP < 0) return fail(KX_EINVAL, "negative counts");
  for (int k = 0; k < nC; k++) {
    int a = w->c_a[k], b = w->c_b[k], t = w->c_type[k];
    if (a < 0 || a >= nB || b >= nB || b < -1)
      return fail(KX_EINVAL, "constraint %d: body index %d/%d outside %d (would be an out-of-bounds read)", k, a, b, nB);
    if (t < KX_PIN || t > KX_GRIP) return fail(KX_EINVAL, "constraint %d: type %d", k, t);
    if (b == a) return fail(KX_EINVAL, "constraint %d couples body %d with itself", k, a);
  }
  h->nB = nB; h->nC = nC; h->nP = nP; h->nR = nR;
  h->prm = *p;
  h->path = p->path;
  // colouring + stable sort (handle order)
  std::vector<int> col;
  TRY(colour_constraints(nB, nC, w->c_a, w->c_b, w->c_color, col));
  h->ord.resize(nC);
  std::iota(h->ord.begin(), h->ord.end(), 0);
  std::stable_sort(h->ord.begin(), h->ord.end(), [&](int x, int y) { return col[x] < col[y]; });
  h->colour.resize(nC);
  for (int k = 0; k < nC; k++) h->colour[k] = col[h->ord[k]];
  for (int c = 0; c < KX_MAX_COLORS; c++) h->colorCnt[c] = 0;
  for (int k = 0; k < nC; k++) h->colorCnt[h->colour[k]]++;
  for (int c = 0, acc = 0; c < KX_MAX_COLORS; c++) { h->colorOff[c] = acc; acc += h->colorCnt[c]; }
  // points grouped by body
  std::vector<int32_t> porder(nP), prange(2 * (size_t)nB);
  TRY(kx_group_points(nP, w->p_body, nB, porder.data(), prange.data()));
  // creatures: connected components over bodies
  std::vector<int> par(nB);
  std::iota(par.begin(), par.end(), 0);
  for (int k = 0; k < nC; k++)
    if (w->c_b[k] >= 0) {
      int x = uf_find(par, w->c_a[k]), y = uf_find(par, w->c_b[k]);
      if (x != y) par[std::max(x, y)] = std::min(x, y);
    }
  std::vector<int> comp_of(nB), comp_id(nB, -1);
  int ncomp = 0;
  for (int b = 0; b < nB; b++) {
    int r = uf_find(par, b);
    if (comp_id[r] < 0) comp_id[r] = ncomp++;
    comp_of[b] = comp_id[r];
  }
  h->nCr = ncomp;
  std::vector<int> cmin(ncomp, nB), cmax(ncomp, -1), ccount(ncomp, 0);
  for (int b = 0; b < nB; b++) {
    int c = comp_of[b];
    cmin[c] = std::min(cmin[c], b); cmax[c] = std::max(cmax[c], b); ccount[c]++;
  }
  h->persist_ok = true;
  for (int c = 0; c < ncomp && h->persist_ok; c++) {
    if (cmax[c] - cmin[c] + 1 != ccount[c]) { h->persist_ok = false; h->persist_why = "a creature's bodies are not contiguous"; }
    if (ccount[c] > 255) { h->persist_ok = false; h->persist_why = "a creature has more than 255 bodies"; }
  }
  // internal constraint order: creature-major, handle order inside
  std::vector<int> ccnt(ncomp, 0);
  for (int k = 0; k < nC; k++) ccnt[comp_of[w->c_a[h->ord[k]]]]++;
  h->cr.assign(ncomp, KxCreature{});
  int acc = 0;
  for (int c = 0; c < ncomp; c++) {
    KxCreature &K = h->cr[c];
    K.body_off = cmin[c]; K.nb = ccount[c];
    K.cons_off = acc; K.nc = ccnt[c]; acc += ccnt[c];
    K.pt_off = prange[cmin[c]];
    int np = 0;
    for (int b = cmin[c]; b <= cmax[c]; b++) np += prange[nB + b];
    K.np = np;
    // round slots pack a constraint index in 8 bits, point ranges in 16
    if (K.nc > 255 || K.np > 65535) { h->persist_ok = false; h->persist_why = "creature too large (> 255 constraints)"; }
  }
  h->h2i.resize(nC);
  std::vector<int> fill(ncomp, 0), i2h(nC);
  for (int k = 0; k < nC; k++) {
    int c = comp_of[w->c_a[h->ord[k]]];
    int i = h->cr[c].cons_off + fill[c]++;
    h->h2i[k] = i; i2h[i] = k;
  }
  // per-creature colour offsets (non-empty colours, ascending)
  for (int c = 0; c < ncomp; c++) {
    KxCreature &K = h->cr[c];
    int cnts[KX_MAX_COLORS] = {};
    for (int j = 0; j < K.nc; j++) cnts[h->colour[i2h[K.cons_off + j]]]++;
    int nco = 0, o = 0;
    for (int col2 = 0; col2 < KX_MAX_COLORS; col2++) {
      if (!cnts[col2]) continue;
      K.coff[nco++] = (uint16_t)o;
      o += cnts[col2];
    }
    K.coff[nco] = (uint16_t)o;
    for (int r = nco + 1; r <= KX_MAX_COLORS; r++) K.coff[r] = (uint16_t)o;
    K.ncol = nco;
    h->maxb = std::max(h->maxb, K.nb); h->maxc = std::max(h->maxc, K.nc); h->maxp = std::max(h->maxp, K.np);
  }
  // creature desc → arms, motor slots
  std::vector<int> cslot(nC, -1), carm(nC, -1), arm_tip, arm_grip;
  // per arm, for the state readout: the mantle and the segment chain its
  // motors run along (shoulder motor mantle -> seg0, then seg i -> seg i+1,
  // kraken.js:103-135)
  std::vector<int> arm_mantle, arm_seg_off(1, 0), arm_segs;
  std::vector<int> d2i(nC);
  for (int k = 0; k < nC; k++) d2i[h->ord[k]] = h->h2i[k];
  if (cd && cd->n_creatures > 0) {
    if (cd->n_creatures != ncomp)
      return fail(KX_EINVAL, "creature desc has %d creatures but the constraint graph has %d components", cd->n_creatures, ncomp);
    for (int c = 0; c < ncomp; c++) {
      int a0 = cd->creature_arm_off[c], a1 = cd->creature_arm_off[c + 1];
      if (a1 - a0 > KX_MAX_ARMS || a1 < a0) return fail(KX_EINVAL, "creature %d: %d arms (max %d)", c, a1 - a0, KX_MAX_ARMS);
      h->cr[c].arm_off = (int)arm_tip.size();
      h->cr[c].narms = a1 - a0;
      for (int a = a0; a < a1; a++) {
        int tip = cd->arm_tip[a], g = cd->arm_grip[a];
        if (tip < 0 || tip >= nB || comp_of[tip] != c) return fail(KX_EINVAL, "arm %d: tip body %d not in creature %d", a, tip, c);
        if (g < 0 || g >= nC || w->c_type[g] != KX_GRIP || w->c_a[g] != tip)
          return fail(KX_EINVAL, "arm %d: constraint %d is not the grip of its tip", a, g);
        arm_tip.push_back(tip);
        arm_grip.push_back(d2i[g]);
        carm[d2i[g]] = a - a0;
        int m0 = cd->arm_motor_off[a], m1 = cd->arm_motor_off[a + 1];
        if (m1 - m0 > KX_SLOTS_PER_ARM) return fail(KX_EINVAL, "arm %d has %d motors (max %d)", a, m1 - m0, KX_SLOTS_PER_ARM);
        for (int m = m0; m < m1; m++) {
          int k = cd->motors[m];
          if (k < 0 || k >= nC || w->c_type[k] != KX_ANGLE || comp_of[w->c_a[k]] != c)
            return fail(KX_EINVAL, "arm %d: motor %d is not an angle constraint of creature %d", a, k, c);
          cslot[d2i[k]] = (a - a0) * KX_SLOTS_PER_ARM + (m - m0);
        }
        bool chain = m1 > m0;
        for (int m = m0; m < m1 && chain; m++) {
          const int k = cd->motors[m];
          chain = w->c_b[k] >= 0 && (m == m0 || w->c_a[k] == w->c_b[cd->motors[m - 1]]);
        }
        chain = chain && w->c_b[cd->motors[m1 - 1]] == tip;
        if (!chain && h->state_why.empty())
          h->state_why = "arm " + std::to_string(a) + ": its motors do not run mantle -> segments -> tip";
        arm_mantle.push_back(chain ? w->c_a[cd->motors[m0]] : tip);
        for (int m = m0; chain && m < m1; m++) arm_segs.push_back(w->c_b[cd->motors[m]]);
        if (!chain) arm_segs.push_back(tip);
        arm_seg_off.push_back((int)arm_segs.size());
      }
    }
    h->has_creatures = true;
  }
  h->nArms = (int)arm_tip.size();
  // ── device buffers ──
  KxDev &d = h->d;
  float *f;
  TRY(dalloc(h, &d.px, nB)); TRY(dalloc(h, &d.py, nB)); TRY(dalloc(h, &d.ang, nB));
  TRY(dalloc(h, &d.vx, nB)); TRY(dalloc(h, &d.vy, nB)); TRY(dalloc(h, &d.av, nB));
  TRY(dalloc(h, &d.ppx, nB)); TRY(dalloc(h, &d.ppy, nB)); TRY(dalloc(h, &d.angP, nB));
  TRY(dalloc(h, &d.fx, nB)); TRY(dalloc(h, &d.fy, nB)); TRY(dalloc(h, &d.tq, nB));
  TRY(dalloc(h, &f, nB)); TRY(dput(f, w->inv_mass, nB)); d.im = f;
  TRY(dalloc(h, &f, nB)); TRY(dput(f, w->inv_inertia, nB)); d.ii = f;
  // constraint statics in internal order
  std::vector<int> ia(nC), ib(nC), it(nC), icr(nC);
  std::vector<float> flax(nC), flay(nC), flbx(nC), flby(nC), frest(nC), fmin(nC), fmax(nC);
  for (int k = 0; k < nC; k++) {
    int s = h->ord[k], i = h->h2i[k];
    ia[i] = w->c_a[s]; ib[i] = w->c_b[s]; it[i] = w->c_type[s]; icr[i] = comp_of[w->c_a[s]];
    flax[i] = w->c_lax[s]; flay[i] = w->c_lay[s]; flbx[i] = w->c_lbx[s]; flby[i] = w->c_lby[s];
    frest[i] = w->c_rest[s]; fmin[i] = w->c_min[s]; fmax[i] = w->c_max[s];
  }
  h->ctype_i = it;
  int *ip;
  TRY(dalloc(h, &ip, nC)); TRY(dput(ip, ia.data(), nC)); d.ca = ip;
  TRY(dalloc(h, &ip, nC)); TRY(dput(ip, ib.data(), nC)); d.cb = ip;
  TRY(dalloc(h, &ip, nC)); TRY(dput(ip, it.data(), nC)); d.ctype = ip;
  TRY(dalloc(h, &ip, nC)); TRY(dput(ip, icr.data(), nC)); d.ccr = ip;
  TRY(dalloc(h, &ip, nC)); TRY(dput(ip, cslot.data(), nC)); d.cslot = ip;
  TRY(dalloc(h, &ip, nC)); TRY(dput(ip, carm.data(), nC)); d.carm = ip;
  TRY(dalloc(h, &f, nC)); TRY(dput(f, flax.data(), nC)); d.lax = f;
  TRY(dalloc(h, &f, nC)); TRY(dput(f, flay.data(), nC)); d.lay = f;
  TRY(dalloc(h, &f, nC)); TRY(dput(f, frest.data(), nC)); d.rest = f;
  TRY(dalloc(h, &f, nC)); TRY(dput(f, fmin.data(), nC)); d.cmin = f;
  TRY(dalloc(h, &f, nC)); TRY(dput(f, fmax.data(), nC)); d.cmax = f;
  TRY(dalloc(h, &d.lbx, nC)); TRY(dput(d.lbx, flbx.data(), nC));
  TRY(dalloc(h, &d.lby, nC)); TRY(dput(d.lby, flby.data(), nC));
  TRY(dalloc(h, &d.comp, nC)); TRY(dalloc(h, &d.target, nC)); TRY(dalloc(h, &d.bf, nC));
  {
    std::vector<float> big(nC, 1e30f);
    TRY(dput(d.bf, big.data(), nC));
  }
  TRY(dalloc(h, &d.broken, nC));
  TRY(dalloc(h, &d.lam, nC)); TRY(dalloc(h, &d.lamx, nC)); TRY(dalloc(h, &d.lamy, nC));
  TRY(dalloc(h, &ip, nC)); TRY(dput(ip, h->h2i.data(), nC)); d.col_list = ip;
  // points
  std::vector<float> plx(nP), ply(nP), pr(nP);
  float rmax = 0.0f, reach = 0.0f;
  for (int q = 0; q < nP; q++) {
    int s = porder[q];
    plx[q] = w->p_lx[s]; ply[q] = w->p_ly[s]; pr[q] = w->p_r[s];
    rmax = std::max(rmax, std::fabs(w->p_r[s]));
    reach = std::max(reach, std::sqrt(w->p_lx[s] * w->p_lx[s] + w->p_ly[s] * w->p_ly[s]));
  }
  TRY(dalloc(h, &f, nP)); TRY(dput(f, plx.data(), nP)); d.plx = f;
  TRY(dalloc(h, &f, nP)); TRY(dput(f, ply.data(), nP)); d.ply = f;
  TRY(dalloc(h, &f, nP)); TRY(dput(f, pr.data(), nP)); d.pr = f;
  TRY(dalloc(h, &ip, nB)); TRY(dput(ip, prange.data(), nB)); d.bpoff = ip;
  TRY(dalloc(h, &ip, nB)); TRY(dput(ip, prange.data() + nB, nB)); d.bpcnt = ip;
  // map
  float4 *r4;
  TRY(dalloc(h, &r4, nR));
  TRY(dput((float *)r4, w->rects, 4 * (size_t)nR));
  d.rects = r4; d.nr = nR;
  TRY(build_grid(h, w->rects, nR, rmax, reach));
  // morphology templates (creature-local static structure, deduplicated)
  if (h->persist_ok) {
    std::map<std::string, int> seen;
    h->cr_tmpl.assign(ncomp, 0);
    std::vector<int> arm_of_grip(nC, -1);
    for (int k = 0; k < nC; k++) if (carm[k] >= 0) arm_of_grip[k] = carm[k];
    for (int c = 0; c < ncomp; c++) {
      const KxCreature &K = h->cr[c];
      TmplHost t;
      t.nb = K.nb;
      for (int j = 0; j < K.nc; j++) {
        int i = K.cons_off + j;
        t.type.push_back(it[i]); t.a.push_back(ia[i] - K.body_off); t.b.push_back(ib[i] < 0 ? -1 : ib[i] - K.body_off);
        t.lax.push_back(flax[i]); t.lay.push_back(flay[i]); t.rest.push_back(frest[i]);
        t.mn.push_back(fmin[i]); t.mx.push_back(fmax[i]);
        if (cslot[i] >= 0) { t.mot_k.push_back(j); t.mot_slot.push_back(cslot[i]); }
        if (it[i] == KX_GRIP) { t.grip_k.push_back(j); t.grip_arm.push_back(arm_of_grip[i]); }
      }
      for (int b = 0; b < K.nb; b++) {
        int g = K.body_off + b;
        t.im.push_back(w->inv_mass[g]); t.ii.push_back(w->inv_inertia[g]);
        t.poff.push_back(prange[g] - K.pt_off); t.pcnt.push_back(prange[nB + g]);
      }
      for (int q = K.pt_off; q < K.pt_off + K.np; q++) { t.plx.push_back(plx[q]); t.ply.push_back(ply[q]); t.pr.push_back(pr[q]); }
      for (int a = 0; a < K.narms; a++) {
        t.arm_tip.push_back(arm_tip[K.arm_off + a] - K.body_off);
        t.arm_grip.push_back(arm_grip[K.arm_off + a] - K.cons_off);
      }
      std::string key = kx_template_key(t);
      auto f = seen.find(key);
      if (f == seen.end()) {
        int id = (int)h->tmpl_host.size();
        seen.emplace(key, id);
        h->tmpl_host.push_back(std::move(t));
        h->cr_tmpl[c] = id;
      } else {
        h->cr_tmpl[c] = f->second;
      }
    }
    if (p->lanes_per_creature) h->G = p->lanes_per_creature;
    if (p->waves_per_block) h->wpb = p->waves_per_block;
    if (h->G != 8 && h->G != 16 && h->G != 32 && h->G != 64) return fail(KX_EINVAL, "lanes_per_creature %d", h->G);
    if (h->wpb != 1 && h->wpb != 2 && h->wpb != 4) return fail(KX_EINVAL, "waves_per_block %d", h->wpb);
    const int r = build_persist(h);
    if (r < 0) {
      // a world the persistent shape cannot hold still runs the plain path;
      // asking for the persistent path fails with the reason
      if (h->path == KX_PATH_PERSISTENT) return r;
      h->persist_ok = false;
      h->persist_why = g_err;
    }
  }
  // creatures / arms
  KxCreature *dcr;
  TRY(dalloc(h, &dcr, ncomp)); TRY(dput(dcr, h->cr.data(), ncomp)); d.cr = dcr;
  TRY(dalloc(h, &ip, arm_tip.size())); TRY(dput(ip, arm_tip.data(), arm_tip.size())); d.arm_tip = ip;
  TRY(dalloc(h, &ip, arm_grip.size())); TRY(dput(ip, arm_grip.data(), arm_grip.size())); d.arm_grip = ip;
  TRY(dalloc(h, &ip, arm_mantle.size())); TRY(dput(ip, arm_mantle.data(), arm_mantle.size())); d.arm_mantle = ip;
  TRY(dalloc(h, &ip, arm_seg_off.size())); TRY(dput(ip, arm_seg_off.data(), arm_seg_off.size())); d.arm_seg_off = ip;
  TRY(dalloc(h, &ip, arm_segs.size())); TRY(dput(ip, arm_segs.data(), arm_segs.size())); d.arm_segs = ip;
  {
    std::vector<float> zero((size_t)ncomp * 2, 0.0f);
    TRY(dalloc(h, &d.prev_v, zero.size())); TRY(dput(d.prev_v, zero.data(), zero.size()));
  }
  TRY(dalloc(h, &d.cd, (size_t)ncomp * KX_MAX_ARMS));
  TRY(dalloc(h, &d.org, (size_t)ncomp * 2));
  TRY(dalloc(h, &ip, nB)); TRY(dput(ip, comp_of.data(), nB)); d.bcr = ip;
  TRY(dalloc(h, &h->step_base, 1));
  // constants (params uniform, gpuworld.js:212-223)
  d.h = (float)p->h;
  d.invH = (float)(1.0 / p->h);
  d.invH2 = (float)(1.0 / (p->h * p->h));
  d.linDamp = (float)std::exp(-p->lin_drag * p->h);
  d.angDamp = (float)std::exp(-p->ang_drag * p->h);
  d.maxSpeed = (float)p->max_speed;
  d.maxSpeed2 = speed_threshold(d.maxSpeed);
  d.maxAngSpeed = (float)p->max_ang_speed;
  d.hd = p->h;
  d.S = p->substeps; d.I = p->iterations;
  d.nB = nB; d.nC = nC; d.nCr = ncomp; d.nArms = h->nArms;
  d.ncol = 0;
  for (int c = 0; c < KX_MAX_COLORS; c++)
    if (h->colorCnt[c]) { d.col_off[d.ncol] = h->colorOff[c]; d.col_cnt[d.ncol] = h->colorCnt[c]; d.ncol++; }
  d.use_grid = 0;
  d.events = nullptr;
  HIPCHK(hipGetDevice(&h->dev));
  HIPCHK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&h->e0));
  HIPCHK(hipEventCreate(&h->e1));
  if (h->path == KX_PATH_PERSISTENT && !h->persist_ok)
    return fail(KX_ENOTSUP, "persistent path unavailable: %s", h->persist_why.c_str());
  return 0;
}

extern "C" int kx_create(const kx_world_desc *desc, const kx_creature_desc *creatures, const kx_params *params,
                         kx_handle **out) {
  if (!desc || !params || !out) return fail(KX_EINVAL, "null argument");
  *out = nullptr;
  kx_handle *h = new (std::nothrow) kx_handle();
  if (!h) return fail(KX_ENOMEM, "out of host memory");
  int r = create_impl(desc, creatures, params, h);
  if (r) { std::string keep = g_err; destroy_impl(h); g_err = keep; return r; }
  *out = h;
  return 0;
}

extern "C" int kx_counts(const kx_handle *h, int32_t *nb, int32_t *nc, int32_t *np, int32_t *ncr) {
  if (!h) return fail(KX_EINVAL, "null handle");
  if (nb) *nb = h->nB;
  if (nc) *nc = h->nC;
  if (np) *np = h->nP;
  if (ncr) *ncr = h->nCr;
  return 0;
}

extern "C" int kx_constraint_order(const kx_handle *h, int32_t *order, int32_t *color) {
  if (!h) return fail(KX_EINVAL, "null handle");
  for (int k = 0; k < h->nC; k++) {
    if (order) order[k] = h->ord[k];
    if (color) color[k] = h->colour[k];
  }
  return 0;
}

extern "C" int kx_color_info(const kx_handle *h, uint32_t *off, uint32_t *cnt) {
  if (!h) return fail(KX_EINVAL, "null handle");
  for (int c = 0; c < KX_MAX_COLORS; c++) { off[c] = h->colorOff[c]; cnt[c] = h->colorCnt[c]; }
  return 0;
}

// ── state transfer ──────────────────────────────────────────────────
static int put_body(kx_handle *h, float *dst, const float *src, const float *fallback) {
  const float *s = src ? src : fallback;
  if (s) return dput(dst, s, h->nB);
  return 0;
}
static int put_cons_f(kx_handle *h, float *dst, const float *src, bool bforce) {
  if (!src) return 0;
  h->hf.resize(h->nC);
  for (int k = 0; k < h->nC; k++) {
    float v = src[k];
    if (bforce) v = std::isfinite(v) ? std::min(v, 1e30f) : 1e30f;   // gpuworld.js:278
    h->hf[h->h2i[k]] = v;
  }
  return dput(dst, h->hf.data(), h->nC);
}
static int put_cons_u8(kx_handle *h, int *dst, const uint8_t *src) {
  if (!src) return 0;
  h->hi.resize(h->nC);
  for (int k = 0; k < h->nC; k++) h->hi[h->h2i[k]] = src[k] ? 1 : 0;
  return dput(dst, h->hi.data(), h->nC);
}

extern "C" int kx_upload(kx_handle *h, const kx_body_state *b, const kx_cons_state *c) {
  if (!h) return fail(KX_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  if (b) {
    if (!b->px || !b->py || !b->ang || !b->vx || !b->vy || !b->av) return fail(KX_EINVAL, "upload needs px,py,ang,vx,vy,av");
    TRY(put_body(h, h->d.px, b->px, nullptr)); TRY(put_body(h, h->d.py, b->py, nullptr));
    TRY(put_body(h, h->d.ang, b->ang, nullptr)); TRY(put_body(h, h->d.vx, b->vx, nullptr));
    TRY(put_body(h, h->d.vy, b->vy, nullptr)); TRY(put_body(h, h->d.av, b->av, nullptr));
    TRY(put_body(h, h->d.ppx, b->ppx, b->px)); TRY(put_body(h, h->d.ppy, b->ppy, b->py));
    TRY(put_body(h, h->d.angP, b->angP, b->ang));
    bool any = false;
    for (float *fa : {b->fx, b->fy, b->tq})
      if (fa) for (int i = 0; i < h->nB && !any; i++) any = fa[i] != 0.0f;
    if (b->fx) TRY(put_body(h, h->d.fx, b->fx, nullptr));
    if (b->fy) TRY(put_body(h, h->d.fy, b->fy, nullptr));
    if (b->tq) TRY(put_body(h, h->d.tq, b->tq, nullptr));
    h->forces_pending = any;
  }
  if (c) {
    TRY(put_cons_f(h, h->d.lbx, c->lbx, false)); TRY(put_cons_f(h, h->d.lby, c->lby, false));
    TRY(put_cons_f(h, h->d.comp, c->compliance, false)); TRY(put_cons_f(h, h->d.target, c->target, false));
    TRY(put_cons_f(h, h->d.bf, c->break_force, true)); TRY(put_cons_u8(h, h->d.broken, c->broken));
    TRY(put_cons_f(h, h->d.lam, c->lambda, false)); TRY(put_cons_f(h, h->d.lamx, c->lambda_x, false));
    TRY(put_cons_f(h, h->d.lamy, c->lambda_y, false));
  }
  return 0;
}

extern "C" int kx_pack_controls(kx_handle *h, const kx_cons_state *c) {
  if (!h || !c) return fail(KX_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(h->st));
  TRY(put_cons_f(h, h->d.lbx, c->lbx, false)); TRY(put_cons_f(h, h->d.lby, c->lby, false));
  TRY(put_cons_f(h, h->d.comp, c->compliance, false)); TRY(put_cons_f(h, h->d.target, c->target, false));
  TRY(put_cons_f(h, h->d.bf, c->break_force, true)); TRY(put_cons_u8(h, h->d.broken, c->broken));
  return 0;
}

static int get_f(kx_handle *h, float *dst, const float *src, int n) {
  if (dst) HIPCHK(hipMemcpy(dst, src, (size_t)n * 4, hipMemcpyDeviceToHost));
  return 0;
}
static int get_cons_f(kx_handle *h, float *dst, const float *src) {
  if (!dst) return 0;
  h->hf.resize(h->nC);
  HIPCHK(hipMemcpy(h->hf.data(), src, (size_t)h->nC * 4, hipMemcpyDeviceToHost));
  for (int k = 0; k < h->nC; k++) dst[k] = h->hf[h->h2i[k]];
  return 0;
}

extern "C" int kx_download(kx_handle *h, kx_body_state *b, kx_cons_state *c, kx_stats *s) {
  if (!h) return fail(KX_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  std::vector<float> tvx, tvy, tav, tpx, tpy, tang;
  if (b) {
    TRY(get_f(h, b->px, h->d.px, h->nB)); TRY(get_f(h, b->py, h->d.py, h->nB));
    TRY(get_f(h, b->vx, h->d.vx, h->nB)); TRY(get_f(h, b->vy, h->d.vy, h->nB));
    TRY(get_f(h, b->ang, h->d.ang, h->nB)); TRY(get_f(h, b->av, h->d.av, h->nB));
    TRY(get_f(h, b->ppx, h->d.ppx, h->nB)); TRY(get_f(h, b->ppy, h->d.ppy, h->nB));
    TRY(get_f(h, b->angP, h->d.angP, h->nB));
    TRY(get_f(h, b->fx, h->d.fx, h->nB)); TRY(get_f(h, b->fy, h->d.fy, h->nB)); TRY(get_f(h, b->tq, h->d.tq, h->nB));
  }
  if (s) {
    const float *vx = b && b->vx ? b->vx : nullptr, *vy = b && b->vy ? b->vy : nullptr;
    const float *av = b && b->av ? b->av : nullptr, *px = b && b->px ? b->px : nullptr;
    const float *py = b && b->py ? b->py : nullptr, *an = b && b->ang ? b->ang : nullptr;
    auto need = [&](const float *have, std::vector<float> &tmp, const float *dev) -> const float * {
      if (have) return have;
      tmp.resize(h->nB);
      (void)hipMemcpy(tmp.data(), dev, (size_t)h->nB * 4, hipMemcpyDeviceToHost);
      return tmp.data();
    };
    vx = need(vx, tvx, h->d.vx); vy = need(vy, tvy, h->d.vy); av = need(av, tav, h->d.av);
    px = need(px, tpx, h->d.px); py = need(py, tpy, h->d.py); an = need(an, tang, h->d.ang);
    // mirrorInto statistics (gpuworld.js:425-441)
    double maxS2 = 0, maxA = 0;
    int64_t nan = 0;
    for (int i = 0; i < h->nB; i++) {
      double x = vx[i], y = vy[i];
      double s2 = x * x + y * y;
      if (s2 > maxS2) maxS2 = s2;
      double aa = std::fabs((double)av[i]);
      if (aa > maxA) maxA = aa;
      if (!std::isfinite(px[i]) || !std::isfinite(py[i]) || !std::isfinite(vx[i]) || !std::isfinite(vy[i]) ||
          !std::isfinite(an[i]))
        nan++;
    }
    s->max_speed = std::sqrt(maxS2); s->max_ang_speed = maxA; s->nan = nan;
  }
  if (c) {
    TRY(get_cons_f(h, c->lbx, h->d.lbx)); TRY(get_cons_f(h, c->lby, h->d.lby));
    TRY(get_cons_f(h, c->compliance, h->d.comp)); TRY(get_cons_f(h, c->target, h->d.target));
    TRY(get_cons_f(h, c->break_force, h->d.bf));
    TRY(get_cons_f(h, c->lambda, h->d.lam)); TRY(get_cons_f(h, c->lambda_x, h->d.lamx));
    TRY(get_cons_f(h, c->lambda_y, h->d.lamy));
    if (c->broken) {
      h->hi.resize(h->nC);
      HIPCHK(hipMemcpy(h->hi.data(), h->d.broken, (size_t)h->nC * 4, hipMemcpyDeviceToHost));
      for (int k = 0; k < h->nC; k++) c->broken[k] = h->hi[h->h2i[k]] ? 1 : 0;
    }
  }
  return 0;
}

extern "C" int kx_get_cooldown(kx_handle *h, double *cd) {
  if (!h || !cd) return fail(KX_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipMemcpy(cd, h->d.cd, (size_t)h->nCr * KX_MAX_ARMS * 8, hipMemcpyDeviceToHost));
  return 0;
}
extern "C" int kx_set_cooldown(kx_handle *h, const double *cd) {
  if (!h || !cd) return fail(KX_EINVAL, "null argument");
  HIPCHK(hipStreamSynchronize(h->st));
  HIPCHK(hipMemcpy(h->d.cd, cd, (size_t)h->nCr * KX_MAX_ARMS * 8, hipMemcpyHostToDevice));
  return 0;
}

// ── stepping ────────────────────────────────────────────────────────
static KxLaw to_law(const kx_law *l) {
  KxLaw L{};
  L.seed = l->seed;
  double t = std::max(0.0, std::min(1.0, l->tone));
  L.comp = 1.0 / (l->stiff_min + t * (l->stiff_max - l->stiff_min));
  L.shoulder_limit = l->shoulder_limit; L.joint_limit = l->joint_limit;
  L.grip_min = l->grip_min; L.grip_max = l->grip_max; L.slip_max = l->slip_max;
  L.cooldown = l->cooldown; L.attach_min = l->attach_min; L.release_below = l->release_below;
  L.stiff_min = l->stiff_min; L.stiff_max = l->stiff_max;
  L.cbase = l->creature_base; L.flags = l->flags;
  return L;
}

static int run_steps(kx_handle *h, const KxLaw *law, int step0, const int *step_base, int nsteps, uint32_t *events) {
  KxDev d = h->d;
  d.events = events;
  int launches = 0;
  int with_forces = h->forces_pending ? 1 : 0;
  h->forces_pending = false;
  if (h->path == KX_PATH_PERSISTENT) {
    int r = kx_persist_launch(h->launch_G, h->wpb, d, law, step0, step_base, nsteps, 0, with_forces, h->pc, h->st);
    if (r) return fail(KX_EDEVICE, "persistent launch failed: %s", hipGetErrorString(hipGetLastError()));
    launches = 1;
  } else {
    for (int s = 0; s < nsteps; s++) {
      int r = step_base ? kx_plain_step(d, law, s, step_base, with_forces && s == 0, s, s, h->st, &launches)
                        : kx_plain_step(d, law, step0 + s, nullptr, with_forces && s == 0, s, s, h->st, &launches);
      if (r) return fail(KX_EDEVICE, "plain launch failed: %s", hipGetErrorString(hipGetLastError()));
    }
  }
  h->last_launches = launches;
  return 0;
}

extern "C" int kx_step(kx_handle *h, int32_t nsteps) {
  if (!h) return fail(KX_EINVAL, "null handle");
  if (nsteps < 0) return fail(KX_EINVAL, "nsteps < 0");
  if (nsteps == 0) return 0;
  HIPCHK(hipEventRecord(h->e0, h->st));
  TRY(run_steps(h, nullptr, 0, nullptr, nsteps, nullptr));
  HIPCHK(hipEventRecord(h->e1, h->st));
  h->timed = true;
  return 0;
}

extern "C" int kx_substep(kx_handle *h, int32_t n) {
  if (!h) return fail(KX_EINVAL, "null handle");
  if (n < 0) return fail(KX_EINVAL, "n < 0");
  if (n == 0) return 0;
  const int with_forces = h->forces_pending ? 1 : 0;
  h->forces_pending = false;
  int launches = 0;
  HIPCHK(hipEventRecord(h->e0, h->st));
  if (kx_plain_substeps(h->d, n, with_forces, h->st, &launches) != 0)
    return fail(KX_EDEVICE, "substep launch failed: %s", hipGetErrorString(hipGetLastError()));
  HIPCHK(hipEventRecord(h->e1, h->st));
  h->last_launches = launches;
  h->timed = true;
  return 0;
}

static int rollout_checks(kx_handle *h, const kx_law *law, int32_t step0, int32_t nsteps) {
  if (!h || !law) return fail(KX_EINVAL, "null argument");
  if (!h->has_creatures) return fail(KX_ENOTSUP, "rollout needs a creature description at create");
  if (nsteps < 0 || step0 < 0) return fail(KX_EINVAL, "bad step range");
  if ((law->flags & KX_LAW_TAPE) && nsteps > h->tape_rows)
    return fail(KX_EINVAL, "tape rollout of %d steps with %d tape rows", nsteps, h->tape_rows);
  return 0;
}

// The rollout proper: the steps on the handle's stream with the event words
// written to `ev_dev` (zeroed by the caller, nsteps * creatures words) when
// given, then copied to `events_host`. Allocates nothing.
int kx_rollout_into(kx_handle *h, const kx_law *law, int32_t step0, int32_t nsteps, uint32_t *ev_dev,
                    uint32_t *events_host) {
  TRY(rollout_checks(h, law, step0, nsteps));
  if (nsteps == 0) return 0;
  KxLaw L = to_law(law);
  HIPCHK(hipEventRecord(h->e0, h->st));
  TRY(run_steps(h, &L, step0, nullptr, nsteps, events_host ? ev_dev : nullptr));
  HIPCHK(hipEventRecord(h->e1, h->st));
  h->timed = true;
  if (events_host) {
    HIPCHK(hipMemcpyAsync(events_host, ev_dev, (size_t)nsteps * h->nCr * 4, hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
  }
  return 0;
}

int kx_handle_basics(const kx_handle *h, int *device, int *n_creatures, hipStream_t *st) {
  if (!h) return fail(KX_EINVAL, "null handle");
  *device = h->dev;
  *n_creatures = h->nCr;
  *st = h->st;
  return 0;
}

extern "C" int kx_rollout(kx_handle *h, const kx_law *law, int32_t step0, int32_t nsteps, uint32_t *events_host) {
  TRY(rollout_checks(h, law, step0, nsteps));
  if (nsteps == 0) return 0;
  size_t need = (size_t)nsteps * h->nCr;
  if (events_host) {
    if (need > h->ev_cap) {
      if (h->ev_dev) HIPCHK(hipFree(h->ev_dev));
      h->ev_dev = nullptr;
      HIPCHK(hipMalloc(&h->ev_dev, need * 4));
      h->ev_cap = need;
    }
    HIPCHK(hipMemsetAsync(h->ev_dev, 0, need * 4, h->st));
  }
  return kx_rollout_into(h, law, step0, nsteps, h->ev_dev, events_host);
}

extern "C" int kx_upload_tape(kx_handle *h, const float *tape, int32_t n_rows) {
  if (!h || (n_rows > 0 && !tape)) return fail(KX_EINVAL, "null argument");
  if (n_rows < 0) return fail(KX_EINVAL, "negative row count");
  const size_t n = (size_t)n_rows * h->nCr * KX_TAPE_WORDS;
  HIPCHK(hipStreamSynchronize(h->st));
  if (n > h->tape_cap) {
    if (h->tape_dev) HIPCHK(hipFree(h->tape_dev));
    h->tape_dev = nullptr;
    h->tape_cap = 0;
    HIPCHK(hipMalloc(&h->tape_dev, n * 4));
    h->tape_cap = n;
    h->d.tape = h->tape_dev;
    // a captured rollout holds the old tape address
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; h->g_nsteps = -1; }
  }
  if (n) HIPCHK(hipMemcpyAsync(h->tape_dev, tape, n * 4, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->tape_rows = n_rows;
  return 0;
}

extern "C" int kx_rollout_graph(kx_handle *h, const kx_law *law, int32_t step0, int32_t nsteps, int32_t reps) {
  if (!h || !law) return fail(KX_EINVAL, "null argument");
  if (!h->has_creatures) return fail(KX_ENOTSUP, "rollout needs a creature description at create");
  if (nsteps <= 0 || reps <= 0 || step0 < 0) return fail(KX_EINVAL, "bad step range");
  KxLaw L = to_law(law);
  if ((L.flags & KX_LAW_TAPE) && nsteps > h->tape_rows)
    return fail(KX_EINVAL, "tape rollout of %d steps with %d tape rows", nsteps, h->tape_rows);
  HIPCHK(hipStreamSynchronize(h->st));
  bool rebuild = !h->gexec || h->g_nsteps != nsteps || h->g_path != h->path || memcmp(&h->g_law, &L, sizeof L) != 0;
  // (kx_set_path drops the cached graph when the persistent shape changes)
  if (rebuild) {
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal));
    int r = 0;
    if (h->path == KX_PATH_PERSISTENT) {
      KxDev d = h->d;
      d.events = nullptr;
      r = kx_persist_launch(h->launch_G, h->wpb, d, &L, 0, h->step_base, nsteps, 0, 0, h->pc, h->st);
    } else {
      KxDev d = h->d;
      d.events = nullptr;
      int launches = 0;
      for (int s = 0; s < nsteps && !r; s++) r = kx_plain_step(d, &L, s, h->step_base, 0, 0, s, h->st, &launches);
    }
    if (!r) r = kx_bump_counter(h->step_base, nsteps, h->st);
    hipError_t ec = hipStreamEndCapture(h->st, &g);
    if (r || ec != hipSuccess) return fail(KX_EDEVICE, "graph capture failed: %s", hipGetErrorString(ec));
    HIPCHK(hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    h->g_nsteps = nsteps; h->g_path = h->path; h->g_law = L;
  }
  HIPCHK(hipMemcpy(h->step_base, &step0, 4, hipMemcpyHostToDevice));
  HIPCHK(hipEventRecord(h->e0, h->st));
  for (int r = 0; r < reps; r++) HIPCHK(hipGraphLaunch(h->gexec, h->st));
  HIPCHK(hipEventRecord(h->e1, h->st));
  h->timed = true;
  h->last_launches = reps;
  return 0;
}

extern "C" int kx_sync(kx_handle *h) {
  if (!h) return fail(KX_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  return 0;
}

extern "C" int kx_last_timing(const kx_handle *h, double *ms, int32_t *launches) {
  if (!h) return fail(KX_EINVAL, "null handle");
  if (!h->timed) return fail(KX_EINVAL, "nothing timed yet");
  HIPCHK(hipEventSynchronize(h->e1));
  float t = 0;
  HIPCHK(hipEventElapsedTime(&t, h->e0, h->e1));
  if (ms) *ms = t;
  if (launches) *launches = h->last_launches;
  return 0;
}

extern "C" int kx_set_path(kx_handle *h, int32_t path, int32_t G, int32_t wpb) {
  if (!h) return fail(KX_EINVAL, "null handle");
  if (path != KX_PATH_PLAIN && path != KX_PATH_PERSISTENT) return fail(KX_EINVAL, "unknown path %d", path);
  if (path == KX_PATH_PERSISTENT && !h->persist_ok)
    return fail(KX_ENOTSUP, "persistent path unavailable: %s", h->persist_why.c_str());
  if (G != 0 && G != 8 && G != 16 && G != 32 && G != 64) return fail(KX_EINVAL, "lanes_per_creature %d", G);
  if (wpb != 0 && wpb != 1 && wpb != 2 && wpb != 4) return fail(KX_EINVAL, "waves_per_block %d", wpb);
  HIPCHK(hipStreamSynchronize(h->st));
  h->path = path;
  if (path == KX_PATH_PERSISTENT && ((G && G != h->G) || (wpb && wpb != h->wpb))) {
    int oldG = h->G, oldW = h->wpb;
    if (G) h->G = G;
    if (wpb) h->wpb = wpb;
    int r = build_persist(h);
    if (r) { h->G = oldG; h->wpb = oldW; std::string keep = g_err; build_persist(h); g_err = keep; return r; }
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
  }
  return 0;
}

extern "C" int kx_schedule_info(const kx_handle *h, int32_t *rounds, int32_t max) {
  if (!h) return fail(KX_EINVAL, "null handle");
  int n = (int)h->tmpl_rounds.size();
  for (int i = 0; i < n && i < max; i++) rounds[i] = h->tmpl_rounds[i];
  return n;
}

extern "C" int kx_persist_info(const kx_handle *h, int32_t *info, int32_t max) {
  if (!h || !info) return fail(KX_EINVAL, "null argument");
  if (!h->persist_ok) return fail(KX_ENOTSUP, "persistent path unavailable: %s", h->persist_why.c_str());
  for (int i = 0; i < 13 && i < max; i++) info[i] = h->pinfo[i];
  return 13;
}

extern "C" int kx_template_info(const kx_handle *h, int32_t t, int32_t *out6) {
  if (!h) return fail(KX_EINVAL, "null handle");
  const int n = (int)h->tmpl_host.size();
  if (t < 0) return n;
  if (t >= n || !out6) return fail(KX_EINVAL, "template %d of %d", t, n);
  const TmplHost &T = h->tmpl_host[t];
  int members = 0;
  for (int c = 0; c < h->nCr; c++) members += h->cr_tmpl[c] == t;
  out6[0] = T.nb; out6[1] = (int)T.type.size(); out6[2] = (int)T.plx.size(); out6[3] = members;
  out6[4] = t < (int)h->tmpl_rounds.size() ? h->tmpl_rounds[t] : 0;
  out6[5] = tmpl_lanes_of(h, t);
  return n;
}

extern "C" int kx_creature_templates(const kx_handle *h, int32_t *tmpl_of_creature) {
  if (!h || !tmpl_of_creature) return fail(KX_EINVAL, "null argument");
  for (int c = 0; c < h->nCr; c++) tmpl_of_creature[c] = h->cr_tmpl[c];
  return h->nCr;
}

extern "C" int kx_set_template_lanes(kx_handle *h, int32_t n, const int32_t *lanes) {
  if (!h) return fail(KX_EINVAL, "null handle");
  if (!h->persist_ok) return fail(KX_ENOTSUP, "persistent path unavailable: %s", h->persist_why.c_str());
  if (n != (int)h->tmpl_host.size() || (n && !lanes)) return fail(KX_EINVAL, "%d lane counts for %d templates", n,
                                                                  (int)h->tmpl_host.size());
  for (int t = 0; t < n; t++)
    if (lanes[t] != 0 && lanes[t] != 8 && lanes[t] != 16 && lanes[t] != 32 && lanes[t] != 64)
      return fail(KX_EINVAL, "template %d: lanes_per_creature %d", t, lanes[t]);
  HIPCHK(hipStreamSynchronize(h->st));
  std::vector<int> old = h->tmpl_lanes;
  h->tmpl_lanes.assign(lanes, lanes + n);
  if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; h->g_nsteps = -1; }
  int r = build_persist(h);
  if (r) {
    std::string keep = g_err;
    h->tmpl_lanes = old;
    build_persist(h);
    g_err = keep;
    return r;
  }
  return 0;
}

// the readout buffer at `stride` floats per creature: one allocation, grown
// (never duplicated) when a call asks for a wider stride, reused for narrower
static int state_buffer(kx_handle *h, int stride) {
  void *p = h->state_dev;
  size_t cap = h->state_cap * sizeof(float);
  TRY(regrow(h, &p, &cap, (size_t)h->nCr * stride * sizeof(float)));
  h->state_dev = (float *)p;
  h->state_cap = cap / sizeof(float);
  h->state_stride = stride;
  return 0;
}

extern "C" int kx_read_state(kx_handle *h, float *out, int32_t stride) {
  if (!h || !out) return fail(KX_EINVAL, "null argument");
  if (!h->has_creatures) return fail(KX_ENOTSUP, "state readout needs the creature structure (kx_creature_desc)");
  if (!h->state_why.empty()) return fail(KX_ENOTSUP, "state readout unavailable: %s", h->state_why.c_str());
  int maxa = 0;
  for (const KxCreature &c : h->cr) maxa = std::max(maxa, c.narms);
  const int need = 6 * maxa + 6 + KX_STATE_EXTRA;
  if (stride < need) return fail(KX_EINVAL, "stride %d: the readout needs %d floats per creature", stride, need);
  const size_t n = (size_t)h->nCr * stride;
  TRY(state_buffer(h, stride));
  if (kx_launch_read_state(h->d, h->state_dev, stride, h->st) != 0) return fail(KX_EDEVICE, "state readout launch failed");
  HIPCHK(hipMemcpyAsync(out, h->state_dev, n * sizeof(float), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  return need;
}

extern "C" int kx_respawn(kx_handle *h, int32_t n, const int32_t *creatures, const double *pose) {
  if (!h || (n > 0 && (!creatures || !pose)) || n < 0) return fail(KX_EINVAL, "null argument");
  if (!h->has_creatures) return fail(KX_ENOTSUP, "respawn needs the creature structure (kx_creature_desc)");
  if (!h->state_why.empty()) return fail(KX_ENOTSUP, "respawn unavailable: %s", h->state_why.c_str());
  if (n == 0) return 0;
  std::vector<char> seen(h->nCr, 0);
  for (int i = 0; i < n; i++) {
    const int c = creatures[i];
    if (c < 0 || c >= h->nCr) return fail(KX_EINVAL, "respawn: creature %d out of range [0, %d)", c, h->nCr);
    if (seen[c]) return fail(KX_EINVAL, "respawn: creature %d listed twice", c);
    seen[c] = 1;
  }
  const size_t ids_bytes = ((size_t)n * 4 + 15) / 16 * 16, need = ids_bytes + (size_t)n * 3 * sizeof(double);
  // the staging buffer may still feed an earlier respawn on the stream
  if (h->resp_dev) HIPCHK(hipStreamSynchronize(h->st));
  TRY(regrow(h, &h->resp_dev, &h->resp_cap, need));
  char *q = (char *)h->resp_dev;
  HIPCHK(hipMemcpy(q, creatures, (size_t)n * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(q + ids_bytes, pose, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice));
  if (kx_launch_respawn(h->d, (const int *)q, (const double *)(q + ids_bytes), n, h->st) != 0)
    return fail(KX_EDEVICE, "respawn launch failed");
  return 0;
}

int kx_policy_io(kx_handle *h, int32_t stride, float **state_dev, float **tape_dev, hipStream_t *st, int32_t *n_cr,
                 int32_t *arms, int launch) {
  if (!h || !state_dev || !tape_dev || !st || !n_cr || !arms) return fail(KX_EINVAL, "null argument");
  if (!h->has_creatures) return fail(KX_ENOTSUP, "the policy step needs the creature structure (kx_creature_desc)");
  if (!h->state_why.empty()) return fail(KX_ENOTSUP, "state readout unavailable: %s", h->state_why.c_str());
  const int a0 = h->cr.empty() ? 0 : h->cr[0].narms;
  for (const KxCreature &c : h->cr)
    if (c.narms != a0) return fail(KX_ENOTSUP, "the policy step needs one arm count for every creature (%d and %d)", a0, c.narms);
  const int need = 6 * a0 + 6 + KX_STATE_EXTRA;
  if (stride < need) return fail(KX_EINVAL, "stride %d: the readout needs %d floats per creature", stride, need);
  *n_cr = h->nCr;
  *arms = a0;
  if (!launch) return need;
  TRY(state_buffer(h, stride));
  const size_t row = (size_t)h->nCr * KX_TAPE_WORDS;
  if (row > h->tape_cap) {
    HIPCHK(hipStreamSynchronize(h->st));
    if (h->tape_dev) HIPCHK(hipFree(h->tape_dev));
    h->tape_dev = nullptr;
    h->tape_cap = 0;
    HIPCHK(hipMalloc(&h->tape_dev, row * 4));
    h->tape_cap = row;
    h->d.tape = h->tape_dev;
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; h->g_nsteps = -1; }
  }
  if (h->tape_rows < 1) h->tape_rows = 1;
  if (kx_launch_read_state(h->d, h->state_dev, stride, h->st) != 0) return fail(KX_EDEVICE, "state readout launch failed");
  *state_dev = h->state_dev;
  *tape_dev = h->tape_dev;
  *st = h->st;
  return need;
}

extern "C" int kx_set_register_slots(kx_handle *h, int32_t enable) {
  if (!h) return fail(KX_EINVAL, "null handle");
  if (enable < 0 || enable > 2) return fail(KX_EINVAL, "register slots %d", enable);
  HIPCHK(hipStreamSynchronize(h->st));
  h->reg_mode = enable;
  if (h->persist_ok) {
    TRY(build_persist(h));
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
  }
  return 0;
}

extern "C" int kx_set_collision_fold(kx_handle *h, int32_t fold) {
  if (fold != KX_FOLD_AUTO && fold != KX_FOLD_PARALLEL && fold != KX_FOLD_OWNER)
    return fail(KX_EINVAL, "collision fold %d", fold);
  return kx_set_coll_mode(h, fold);
}

int kx_set_coll_mode(kx_handle *h, int mode) {
  if (!h) return fail(KX_EINVAL, "null handle");
  if (mode < 0 || mode > 2) return fail(KX_EINVAL, "collision mode %d", mode);
  if (mode == h->coll_mode) return 0;
  HIPCHK(hipStreamSynchronize(h->st));
  const int old = h->coll_mode;
  h->coll_mode = mode;
  if (h->persist_ok) {
    int r = build_persist(h);
    if (r) { h->coll_mode = old; std::string keep = g_err; build_persist(h); g_err = keep; return r; }
    if (h->gexec) { (void)hipGraphExecDestroy(h->gexec); h->gexec = nullptr; }
  }
  return 0;
}

int kx_handle_view(kx_handle *h, KxHandleView *v) {
  if (!h || !v) return fail(KX_EINVAL, "null argument");
  if (h->path != KX_PATH_PERSISTENT || !h->persist_ok)
    return fail(KX_ENOTSUP, "batched launches need the persistent path");
  v->d = h->d;
  v->d.events = nullptr;
  v->P = h->pc;
  v->wpb = h->wpb;
  v->st = h->st;
  v->bytes = 0;
  for (auto &b : h->bufs) v->bytes += b.bytes;
  return 0;
}

// plain-path SDF mode switch (0 linear WGSL scan, 1 cull grid) — exported for
// the benchmark's like-for-like comparison
extern "C" int kx_set_plain_sdf(kx_handle *h, int32_t use_grid) {
  if (!h) return fail(KX_EINVAL, "null handle");
  HIPCHK(hipStreamSynchronize(h->st));
  h->d.use_grid = use_grid ? 1 : 0;
  return 0;
}
