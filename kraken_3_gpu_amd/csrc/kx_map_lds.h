// Map queries of the persistent kernels: the culled SDF fold against the
// map staged in LDS, the LDS occupancy bitmap or the map in HBM.
#pragma once
#include "kx_device.h"

// The map staged in LDS: per cell entry the rect itself (ascending rect
// index within a cell, the fold order), then 16-bit cell offsets. Two entries
// are read at once; a cell rarely holds more.
__device__ __forceinline__ bool kx_sdf_cells_lds(const KxDev &d, const uint16_t *coff, const float4 *crect, float x,
                                                 float y, float ox, float oy, KxHit &h) {
  const float gx = (x + ox - d.gx0) * d.inv_cs;
  const float gy = (y + oy - d.gy0) * d.inv_cs;
  if (!(gx >= 0.0f && gy >= 0.0f && gx < (float)d.gnx && gy < (float)d.gny)) return false;
  const int ix = min((int)gx, d.gnx - 1), iy = min((int)gy, d.gny - 1);
  const int cell = iy * d.gnx + ix;
  const int b = coff[cell], e = coff[cell + 1];
  if (b == e) return false;
  const float4 r0 = crect[b], r1 = crect[b + 1 < e ? b + 1 : b];
  kx_hit_init(h);
  kx_fold_rect(x, y, r0, ox, oy, h);
  if (b + 1 < e) kx_fold_rect(x, y, r1, ox, oy, h);
  for (int k = b + 2; k < e; k++) kx_fold_rect(x, y, crect[k], ox, oy, h);
  return isfinite(h.d);
}

// The cull cell of a centre (x, y) relative to the origin (ox, oy): false,
// and cell 0, when the centre lies outside the grid or is not finite.
__device__ __forceinline__ bool kx_centre_cell(const KxDev &d, float x, float y, float ox, float oy, int &cell) {
  const float gx = (x + ox - d.gx0) * d.inv_cs;
  const float gy = (y + oy - d.gy0) * d.inv_cs;
  const bool in = gx >= 0.0f && gy >= 0.0f && gx < (float)d.gnx && gy < (float)d.gny;
  const int ix = min((int)gx, d.gnx - 1), iy = min((int)gy, d.gny - 1);
  cell = in ? iy * d.gnx + ix : 0;
  return in;
}

// Squared distance from (x, y) to the box a rect's clamp maps onto.
__device__ __forceinline__ float kx_box_d2(float x, float y, float4 r, float ox, float oy) {
  const float rx = r.x - ox, ry = r.y - oy;
  const float cx = fminf(fmaxf(x, rx), rx + r.z), cy = fminf(fmaxf(y, ry), ry + r.w);
  const float dx = x - cx, dy = y - cy;
  return dx * dx + dy * dy;
}

// Contact test of a body centre (x, y) (relative to the origin (ox, oy))
// against the map staged in LDS: true when the box of some rect registered in
// the centre's cull cell lies closer than D to the centre. The box of a rect
// is the set its clamp maps onto (kx_fold_rect: min(max(x, rx), rx + w)), a
// convex set, so the distance to it is 1-Lipschitz: a point within reach of
// the centre whose fold finds a rect closer than its radius r has that rect
// within reach + r of the centre. With D at least that and no larger than the
// cull radius, every such rect is registered in the centre's cell (cells hold
// every rect within the cull radius + 1 px). A centre outside the grid lies
// farther than the cull radius + one cell from every rect, and a non-finite
// centre is never near (its points never hit either).
__device__ __forceinline__ bool kx_centre_near_lds(const KxDev &d, const uint16_t *coff, const float4 *crect, float x,
                                                   float y, float ox, float oy, float D) {
  int cell;
  if (!kx_centre_cell(d, x, y, ox, oy, cell)) return false;
  const int b = coff[cell], e = coff[cell + 1];
  const float D2 = D * D;
  bool near = false;
  for (int k = b; k < e; k++) near = near || kx_box_d2(x, y, crect[k], ox, oy) < D2;
  return near;
}

// Grid-culled SDF query against whichever copy of the map the launch uses.
__device__ __forceinline__ bool kx_query(const KxDev &d, const KxPersistCfg &P, const uint32_t *bits,
                                         const uint16_t *m_off, const float4 *m_crect, float wx, float wy,
                                         float ox, float oy, KxHit &hh) {
  return P.map_words ? kx_sdf_cells_lds(d, m_off, m_crect, wx, wy, ox, oy, hh)
         : P.bm_words ? kx_sdf_grid_bm(d, bits, wx, wy, ox, oy, hh)
                      : kx_sdf_grid(d, wx, wy, ox, oy, hh);
}

// Two queries against the map staged in LDS, each exactly kx_sdf_cells_lds:
// both points' cell offsets are read together, then the first two rects of
// both cells (a cell rarely holds more), so two queries cost two LDS round
// trips instead of four. Every index read is valid (0 when unused).
__device__ __forceinline__ void kx_sdf2_cells_lds(const KxDev &d, const uint16_t *coff, const float4 *crect, float x0,
                                                  float y0, float x1, float y1, float ox, float oy, bool &ok0,
                                                  KxHit &h0, bool &ok1, KxHit &h1) {
  auto cellof = [&](float x, float y, int &cell) {
    const float gx = (x + ox - d.gx0) * d.inv_cs;
    const float gy = (y + oy - d.gy0) * d.inv_cs;
    const bool in = gx >= 0.0f && gy >= 0.0f && gx < (float)d.gnx && gy < (float)d.gny;
    const int ix = min((int)gx, d.gnx - 1), iy = min((int)gy, d.gny - 1);
    cell = in ? iy * d.gnx + ix : 0;
    return in;
  };
  int c0, c1;
  const bool in0 = cellof(x0, y0, c0), in1 = cellof(x1, y1, c1);
  const int b0 = coff[c0], e0 = coff[c0 + 1], b1 = coff[c1], e1 = coff[c1 + 1];
  const int n0 = in0 ? e0 - b0 : 0, n1 = in1 ? e1 - b1 : 0;
  const float4 r00 = crect[n0 > 0 ? b0 : 0], r01 = crect[n0 > 1 ? b0 + 1 : 0];
  const float4 r10 = crect[n1 > 0 ? b1 : 0], r11 = crect[n1 > 1 ? b1 + 1 : 0];
  kx_hit_init(h0);
  kx_hit_init(h1);
  if (n0 > 0) kx_fold_rect(x0, y0, r00, ox, oy, h0);
  if (n0 > 1) kx_fold_rect(x0, y0, r01, ox, oy, h0);
  for (int k = 2; k < n0; k++) kx_fold_rect(x0, y0, crect[b0 + k], ox, oy, h0);
  if (n1 > 0) kx_fold_rect(x1, y1, r10, ox, oy, h1);
  if (n1 > 1) kx_fold_rect(x1, y1, r11, ox, oy, h1);
  for (int k = 2; k < n1; k++) kx_fold_rect(x1, y1, crect[b1 + k], ox, oy, h1);
  ok0 = n0 > 0 && isfinite(h0.d);
  ok1 = n1 > 0 && isfinite(h1.d);
}

// The same two queries against the map in HBM (kx_sdf_grid_bm with the
// occupancy bitmap `bits` in LDS, kx_sdf_grid without it): both cells'
// occupancy bits, then both cells' offsets, then the first two rects of both
// cells are read together, two dependent global loads for both points
// instead of two per point. Every index read is valid (0 when unused: the
// offsets have at least two entries and the entry array at least one).
__device__ __forceinline__ void kx_sdf2_cells_glb(const KxDev &d, const uint32_t *bits, float x0, float y0, float x1,
                                                  float y1, float ox, float oy, bool &ok0, KxHit &h0, bool &ok1,
                                                  KxHit &h1) {
  auto cellof = [&](float x, float y, int &cell) {
    const float gx = (x + ox - d.gx0) * d.inv_cs;
    const float gy = (y + oy - d.gy0) * d.inv_cs;
    const bool in = gx >= 0.0f && gy >= 0.0f && gx < (float)d.gnx && gy < (float)d.gny;
    const int ix = min((int)gx, d.gnx - 1), iy = min((int)gy, d.gny - 1);
    cell = in ? iy * d.gnx + ix : 0;
    return in;
  };
  int c0, c1;
  bool in0 = cellof(x0, y0, c0), in1 = cellof(x1, y1, c1);
  if (bits) {
    in0 = in0 && ((bits[c0 >> 5] >> (c0 & 31)) & 1u) != 0u;
    in1 = in1 && ((bits[c1 >> 5] >> (c1 & 31)) & 1u) != 0u;
  }
  const int b0 = d.cell_off[c0], e0 = d.cell_off[c0 + 1], b1 = d.cell_off[c1], e1 = d.cell_off[c1 + 1];
  const int n0 = in0 ? e0 - b0 : 0, n1 = in1 ? e1 - b1 : 0;
  const float4 r00 = d.cell_rects[n0 > 0 ? b0 : 0], r01 = d.cell_rects[n0 > 1 ? b0 + 1 : 0];
  const float4 r10 = d.cell_rects[n1 > 0 ? b1 : 0], r11 = d.cell_rects[n1 > 1 ? b1 + 1 : 0];
  kx_hit_init(h0);
  kx_hit_init(h1);
  if (n0 > 0) kx_fold_rect(x0, y0, r00, ox, oy, h0);
  if (n0 > 1) kx_fold_rect(x0, y0, r01, ox, oy, h0);
  for (int k = 2; k < n0; k++) kx_fold_rect(x0, y0, d.cell_rects[b0 + k], ox, oy, h0);
  if (n1 > 0) kx_fold_rect(x1, y1, r10, ox, oy, h1);
  if (n1 > 1) kx_fold_rect(x1, y1, r11, ox, oy, h1);
  for (int k = 2; k < n1; k++) kx_fold_rect(x1, y1, d.cell_rects[b1 + k], ox, oy, h1);
  ok0 = n0 > 0 && isfinite(h0.d);
  ok1 = n1 > 0 && isfinite(h1.d);
}

