// walls.hpp -- device kernels of the moving-wall (fouling layer) update:
// the fouling-layer node types (updateNType, lsKernels.cl:47-170), the node
// boundary flags (updateNodeBoundaryInfo, :512-580), the interpolated
// bounce-back links and wall spacings (updateIBBOnly / updateBoundaryNodes,
// :221-505) and the wall distance (updateWallD, kOmegaKernels.cl:813-840).
//
// The reference appends boundary links through an atomic counter, so its
// link order (and which q a repeated link keeps) depends on scheduling.
// Here every node searches the wall segments that can cut its links itself,
// in increasing segment order, and keeps the last crossing per direction:
// the same choice the sequential host build makes, with no atomics.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.hpp"

namespace kofm {
namespace dev {

// nType bits (the reference values, ConstDef.h:450-480)
enum : uint32_t {
  kTMSolid = 0x0001,
  kTMFluid = 0x0002,
  kTM0Fluid = 0x0008,
  kTSolidBoundary = 0x0010,
};

struct WallSegments {
  const double* C0;     // 2 * nN: the wall points of the initial (clean) wall
  const double* C;      // 2 * nN: the wall points now
  const int32_t* BL;    // 2 * nBL: point pairs of the boundary links
  int nBL;
  // segments whose x range can reach column i: segIdx[segStart[i] .. segStart[i+1])
  const int32_t* segStart;
  const int32_t* segIdx;
};

struct WallGrid {
  int nX, nY, XF;
  int64_t D;
  double minDx;
};

__device__ __forceinline__ double cross2(double ax, double ay, double bx, double by) {
  return fabs(ax * by - ay * bx);
}

// test_inside (lsKernels.cl:31-41): L0 inside triangle P0 P1 P2 when the
// three sub-triangle areas add up to the whole (to 1e-5).
__device__ __forceinline__ bool insideTriangle(double p0x, double p0y, double p1x, double p1y,
                                               double p2x, double p2y, double lx, double ly) {
  const double v10x = p1x - p0x, v10y = p1y - p0y, v20x = p2x - p0x, v20y = p2y - p0y;
  const double v21x = p2x - p1x, v21y = p2y - p1y, vd0x = lx - p0x, vd0y = ly - p0y;
  const double vd1x = lx - p1x, vd1y = ly - p1y;
  const double st = cross2(v10x, v10y, v20x, v20y), s0 = cross2(v10x, v10y, vd0x, vd0y);
  const double s1 = cross2(vd0x, vd0y, v20x, v20y), s2 = cross2(v21x, v21y, vd1x, vd1y);
  return fabs(st - s0 - s1 - s2) < 1.0e-5;
}

// updateNType (lsKernels.cl:47-170), one work item per boundary link: fluid
// nodes inside the two triangles between the link's clean position
// (C0[a], C0[b]) and its position now (C[a], C[b]) are marked as fouled.
// Several links may mark one node; the mark is a plain store of 1.
__global__ __launch_bounds__(64) void wallFoulMark(WallSegments w, WallGrid g,
                                                   const int16_t* __restrict__ nType,
                                                   uint8_t* __restrict__ mark) {
  const int s = blockIdx.x * 64 + threadIdx.x;
  if (s >= w.nBL) return;
  const int a = w.BL[2 * s], b = w.BL[2 * s + 1];
  const double p0x = w.C0[2 * a], p0y = w.C0[2 * a + 1], p1x = w.C0[2 * b], p1y = w.C0[2 * b + 1];
  const double p2x = w.C[2 * a], p2y = w.C[2 * a + 1], p3x = w.C[2 * b], p3y = w.C[2 * b + 1];
  const bool same02 = p0x == p2x && p0y == p2y, same13 = p1x == p3x && p1y == p3y;
  if (same02 && same13) return;  // no fouling layer on this link
  // bounding box: truncated minimum, ceiled maximum (min4 / max4, Heading.cl:508-527)
  int lox = int(fmin(fmin(p0x, p1x), fmin(p2x, p3x))), loy = int(fmin(fmin(p0y, p1y), fmin(p2y, p3y)));
  int hix = int(ceil(fmax(fmax(p0x, p1x), fmax(p2x, p3x))));
  int hiy = int(ceil(fmax(fmax(p0y, p1y), fmax(p2y, p3y))));
  if (hiy >= g.nY) hiy = g.nY - 1;
  if (loy >= g.nY || hiy < 0) return;
  if (loy < 0) loy = 0;
  if (hix >= g.nX) hix = g.nX - 1;
  if (lox >= g.nX || hix < 0) return;
  if (lox < 0) lox = 0;
  for (int i = lox; i <= hix; i++)
    for (int j = loy; j <= hiy; j++) {
      const int64_t n = i + int64_t(g.XF) * j;
      if (nType[n] & kTMSolid) continue;
      const double lx = double(i), ly = double(j);
      if ((!same13 && insideTriangle(p3x, p3y, p2x, p2y, p1x, p1y, lx, ly)) ||
          (!same02 && insideTriangle(p0x, p0y, p1x, p1y, p2x, p2y, lx, ly)))
        mark[n] = 1;
    }
}

// The marked nodes become solid (SWITCH_FLUID_TO_SOLID keeps M0_FLUID, so a
// fouling-layer node stays recognisable) and get the macros, k, omega and
// eddy viscosity of the reference's update (rho 1, everything else 0).
template <class R>
__global__ __launch_bounds__(kBlockX) void wallFoulApply(WallGrid g, int16_t* __restrict__ nType,
                                                         uint8_t* __restrict__ mark,
                                                         R* __restrict__ rho, R* __restrict__ ux,
                                                         R* __restrict__ uy, double* __restrict__ K,
                                                         double* __restrict__ O,
                                                         double* __restrict__ nut) {
  const int64_t n = int64_t(blockIdx.x) * kBlockX + threadIdx.x;
  if (n >= g.D || !mark[n]) return;
  mark[n] = 0;
  const uint32_t t = uint16_t(nType[n]);
  if (t & kTMSolid) return;
  nType[n] = int16_t(uint16_t(t ^ 0x0003u));
  rho[n] = R(1);
  ux[n] = R(0);
  uy[n] = R(0);
  if (K) {
    K[n] = 0.;
    O[n] = 0.;
    nut[n] = 0.;
  }
}

// updateNodeBoundaryInfo (lsKernels.cl:512-580): the E..NW bound bits of a
// node mark neighbours of the other kind (solid neighbours of a fluid node,
// fluid neighbours of a solid one), x periodic, nothing beyond the bottom
// and top rows; a solid node with a fluid N/E/S/W neighbour is a solid
// boundary node and keeps no direction bits.  The device mask follows
// (solid, solid-boundary and bound bits; the interpolated-link bit is set
// by wallLinks).
__global__ __launch_bounds__(kBlockX) void wallBoundaryInfo(WallGrid g,
                                                            int16_t* __restrict__ nType,
                                                            uint16_t* __restrict__ mask) {
  const int i = blockIdx.x * kBlockX + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nX) return;
  const int64_t n = i + int64_t(g.XF) * j;
  const uint32_t t = uint16_t(nType[n]) & 0x00EFu;  // RESET_NTYPE_NODE
  const uint32_t cmp = (t & 0x3u) ^ 0x3u;
  const int64_t e = i < g.nX - 1 ? n + 1 : int64_t(g.XF) * j;
  const int64_t w = i > 0 ? n - 1 : (g.nX - 1) + int64_t(g.XF) * j;
  auto other = [&](int64_t m) { return (uint32_t(uint16_t(nType[m])) & cmp) != 0; };
  uint32_t mod = 0;
  if (other(e)) mod |= 0x0100u;  // E
  if (other(w)) mod |= 0x0200u;  // W
  if (j < g.nY - 1) {
    if (other(n + g.XF)) mod |= 0x0400u;  // N
    if (other(e + g.XF)) mod |= 0x1000u;  // NE
    if (other(w + g.XF)) mod |= 0x8000u;  // NW
  }
  if (j > 0) {
    if (other(n - g.XF)) mod |= 0x0800u;  // S
    if (other(e - g.XF)) mod |= 0x4000u;  // SE
    if (other(w - g.XF)) mod |= 0x2000u;  // SW
  }
  if (t & kTMSolid) mod = (mod & 0x0F00u) ? kTSolidBoundary : 0u;
  const uint32_t nt = t | mod;
  nType[n] = int16_t(uint16_t(nt));
  uint32_t m = nt & 0xFF00u;
  if (nt & kTMSolid) m |= kSolid;
  if (nt & kTSolidBoundary) m |= kSolidBoundary;
  mask[n] = uint16_t(m);
}

// bcFindIntersectionLS (lsKernels.cl:172-215): the probe from (lx, ly) along
// direction index d (reversed when it points away from the link), the
// fraction of the way to the link line, inside [0, 1] and on the link.
__device__ __forceinline__ bool probeLink(double lx, double ly, int& d, double v0x, double v0y,
                                          double v1x, double v1y, double len, double nx, double ny,
                                          double& dist) {
  double ldx = double(cx(d + 1)), ldy = double(cy(d + 1));
  double den = nx * ldx + ny * ldy;
  if (den == 0.) return false;
  if (den > 0.) {
    d = (d & 1) ? d - 1 : d + 1;  // RevDir
    ldx = -ldx;
    ldy = -ldy;
    den = -den;
  }
  dist = (nx * (v0x - lx) + ny * (v0y - ly)) / den;
  if (dist < 0. || dist > 1.) return false;
  const double cx_ = lx + ldx * dist, cy_ = ly + ldy * dist;
  const double a = sqrt((cx_ - v0x) * (cx_ - v0x) + (cy_ - v0y) * (cy_ - v0y));
  const double b = sqrt((cx_ - v1x) * (cx_ - v1x) + (cy_ - v1y) * (cy_ - v1y));
  return fabs(len - a - b) < 1.0e-5;
}

// Per node: the interpolated bounce-back links of updateIBBOnly
// (lsKernels.cl:429-505, columns 1 .. nX-2; the probes E, N, NE, SE, each
// reversed to face the link, the direction loop continuing from a reversed
// probe) and, for k-omega, the axis wall spacings of updateBoundaryNodes
// (:221-428): a crossing toward a solid neighbour sets this node's spacing
// to max(q, minDx), and the neighbour's opposite spacing to max(1 - q,
// minDx) when it is a fouling-layer node.  Links cut by several segments
// keep the last segment's q.  q8 [8 * D] receives q per direction (-1: no
// link), bits [D] the link directions, mask gains the interpolated-link bit.
__global__ __launch_bounds__(kBlockX) void wallLinks(WallSegments w, WallGrid g,
                                                     const int16_t* __restrict__ nType,
                                                     uint16_t* __restrict__ mask,
                                                     double* __restrict__ q8,
                                                     uint8_t* __restrict__ bits,
                                                     double* __restrict__ dX) {
  const int i = blockIdx.x * kBlockX + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nX) return;
  const int64_t n = i + int64_t(g.XF) * j;
  uint32_t lb = 0;
  double q[8];
#pragma unroll
  for (int d = 0; d < 8; d++) q[d] = -1.;
  const bool fluid = !(uint16_t(nType[n]) & kTMSolid);
  if (fluid) {
    const double lx = double(i), ly = double(j);
    const bool linkCol = i >= 1 && i <= g.nX - 2;
    for (int k = w.segStart[i]; k < w.segStart[i + 1]; k++) {
      const int s = w.segIdx[k];
      const int a = w.BL[2 * s], b = w.BL[2 * s + 1];
      const double v0x = w.C[2 * a], v0y = w.C[2 * a + 1], v1x = w.C[2 * b], v1y = w.C[2 * b + 1];
      // truncated minimum, ceiled maximum (min2 / max2, Heading.cl:493-505)
      const int lox = int(fmin(v0x, v1x)), loy = int(fmin(v0y, v1y));
      const int hix = int(ceil(fmax(v0x, v1x))), hiy = int(ceil(fmax(v0y, v1y)));
      if (i < lox || i > hix || j < loy || j > hiy) continue;
      const double tx = v1x - v0x, ty = v1y - v0y;
      const double len = sqrt(tx * tx + ty * ty);
      const double nx = -(ty / len), ny = tx / len;
      if (linkCol)
        for (int d = 0; d < 8; d += 2) {
          double dist;
          if (probeLink(lx, ly, d, v0x, v0y, v1x, v1y, len, nx, ny, dist)) q[d] = dist;
        }
      if (dX)
        for (int d0 = 0; d0 < 4; d0 += 2) {  // E then N, each possibly reversed
          int d = d0;
          double dist;
          if (!probeLink(lx, ly, d, v0x, v0y, v1x, v1y, len, nx, ny, dist)) continue;
          int si = i + cx(d + 1);
          si = si >= g.nX ? si - g.nX : (si < 0 ? si + g.nX : si);
          const int sj = j + cy(d + 1);
          if (sj < 0 || sj >= g.nY) continue;
          const int64_t m = si + int64_t(g.XF) * sj;
          const uint32_t tm = uint16_t(nType[m]);
          if (!(tm & kTMSolid)) continue;
          dX[n + int64_t(d) * g.D] = fmax(dist, g.minDx);
          if (tm & kTM0Fluid) dX[m + int64_t(d ^ 1) * g.D] = fmax(1. - dist, g.minDx);
        }
    }
  }
#pragma unroll
  for (int d = 0; d < 8; d++) {
    q8[n + int64_t(d) * g.D] = q[d];
    if (q[d] >= 0.) lb |= 1u << d;
  }
  bits[n] = uint8_t(lb);
  if (lb) mask[n] |= kIbbNode;
}

// q rows of the wall nodes (one row of 8 per node in list order) into the
// engine's q table (storage arithmetic type) and an FP64 copy for the host.
template <class R>
__global__ __launch_bounds__(64) void wallLinkGather(const double* __restrict__ q8, int64_t D,
                                                     const int32_t* __restrict__ nodes, int n,
                                                     R* __restrict__ ibbQ,
                                                     double* __restrict__ rows) {
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= n) return;
  const int64_t node = nodes[r];
#pragma unroll
  for (int d = 0; d < 8; d++) {
    const double v = q8[node + int64_t(d) * D];
    ibbQ[8 * int64_t(r) + d] = R(v);
    rows[8 * int64_t(r) + d] = v;
  }
}

// kOmega::updateWallD (kOmegaKernels.cl:813-840) with findMinDist: the
// distance of every fluid node to the nearest link of the bottom range
// [lsMap - rad, lsMap + rad] and of the top range, the link as it is now.
__global__ __launch_bounds__(kBlockX) void wallDistance(WallSegments w, WallGrid g, int rad,
                                                        const uint16_t* __restrict__ lsMap,
                                                        const int16_t* __restrict__ nType,
                                                        double* __restrict__ wallD) {
  const int i = blockIdx.x * kBlockX + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= g.nX) return;
  const int64_t n = i + int64_t(g.XF) * j;
  if (uint16_t(nType[n]) & kTMSolid) {
    wallD[n] = 0.;
    return;
  }
  const int half = w.nBL / 2;
  const int lb = lsMap[i], lt = lsMap[i + g.XF];
  const int b0 = lb - rad > 0 ? lb - rad : 0, b1 = lb + rad < half ? lb + rad : half;
  const int t0 = lt - rad > half ? lt - rad : half, t1 = lt + rad < w.nBL ? lt + rad : w.nBL;
  const double px = double(i), py = double(j);
  double best = 1000.;
  for (int top = 0; top < 2; top++) {
    const int lo = top ? t0 : b0, hi = top ? t1 : b1;
    for (int s = lo; s < hi; s++) {
      const int a = w.BL[2 * s], b = w.BL[2 * s + 1];
      const double P0x = w.C[2 * a], P0y = w.C[2 * a + 1], P1x = w.C[2 * b], P1y = w.C[2 * b + 1];
      double tx = P1x - P0x, ty = P1y - P0y;
      const double tl = sqrt(tx * tx + ty * ty);
      double d = fabs(ty * px - tx * py + P1x * P0y - P1y * P0x) / tl;
      tx = tx / tl;
      ty = ty / tl;
      const double fx = px - (-ty) * d;
      const bool before = top ? fx > P0x : fx < P0x;
      const bool after = top ? fx < P1x : fx > P1x;
      if (before)
        d = sqrt((px - P0x) * (px - P0x) + (py - P0y) * (py - P0y));
      else if (after)
        d = sqrt((px - P1x) * (px - P1x) + (py - P1y) * (py - P1y));
      best = best < d ? best : d;
    }
  }
  wallD[n] = best;
}

}  // namespace dev
}  // namespace kofm
