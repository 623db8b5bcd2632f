// nusselt.hpp -- device kernels of the Nusselt-number output of the
// temperature solve (clVariablesFD::updateTimeData, clVariablesFD.cpp:
// 834-869): the wall-normal nearest node of each column (updateNuCoeff1/2,
// tfdKernels.cl:737-856), the bulk and wall temperatures and the wall
// gradient per column (FD_calc_Nu, getTbm and getTandDT, :858-982) and the
// sums over the columns at the period boundaries of the wavy section
// (getNuMean, :984-1033).  All arithmetic is FP64, as in the reference.
//
// The reference finds the nearest node with one work-item per link and an
// atomic minimum on the distance, then writes the node's row and vector
// outside the atomic, so a tie or two close updates can leave a row and
// vector from different links.  Here one thread per (column, wall) visits
// the links that reach its column in the particle tracer's link order and
// keeps the first strict minimum: the order of the reference's host search
// (clVariablesFD::nuFindNearestNode, clVariablesFD.cpp:498-536).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace kofm {
namespace dev {

// Heading.cl: EPSILON, the tolerance of the foot-in-segment test
constexpr double kNuEpsilon = 1.11e-16;
// no node found: the reference's output value (FD_calc_Nu)
constexpr double kNuNone = -1000.;
// the reference's initial nearest distance (nuDist.fill(100.))
constexpr double kNuFar = 100.;
// getNuMean's work-group: the column is summed by at most this many threads
constexpr int kNuMeanMaxThreads = 1024;

struct NuGeom {
  const double* C;         // 2 * nN: the wall points now (vls.C)
  const int32_t* BL;       // 2 * nBL
  const int32_t* slotStart;  // 2n + 1: links that reach slot s: slotLinks[slotStart[s] ..]
  const int32_t* slotLinks;
  const int16_t* nType;    // D: node types now
  int n;                   // columns: x_stop - x_min
  int xMin, xStop;         // TR_X_IND_START, TR_X_IND_STOP
  int nY, XF;
};

// nuFindIntersectionNormal (tfdKernels.cl:737-756): signed distance from L
// to the line through P0, P1 along its unit normal (-t.y, t.x), and whether
// the foot of the normal lies on the segment (|P1 - P0| = |C - P0| + |C - P1|
// to EPSILON).
__device__ inline bool nuFoot(double Lx, double Ly, double P0x, double P0y, double P1x,
                                       double P1y, double& dist, double& Cx, double& Cy) {
  const double tx = P1x - P0x, ty = P1y - P0y;
  const double nmx = -ty, nmy = tx;
  const double nl = sqrt(nmx * nmx + nmy * nmy);
  const double nx = nmx / nl, ny = nmy / nl;
  dist = (Lx - P0x) * nx + (Ly - P0y) * ny;
  Cx = Lx - nx * dist;
  Cy = Ly - ny * dist;
  const double d0x = Cx - P0x, d0y = Cy - P0y, d1x = Cx - P1x, d1y = Cy - P1y;
  const double l10 = sqrt(tx * tx + ty * ty);
  const double l0 = sqrt(d0x * d0x + d0y * d0y);
  const double l1 = sqrt(d1x * d1x + d1y * d1y);
  return fabs(l10 - l0 - l1) < kNuEpsilon;
}

// updateNuCoeff1/2: slot s < n is column xMin + s of the bottom wall's
// links, slot n + s the same column of the top wall's.  rows[s] =
// (row << 2) | type, type bit 0 when the node lies left of the foot (the
// wall is to its east), bit 1 when it lies below it; row = -1 (and type 0)
// when no node was found.  vec[s] = node - foot.
__global__ void nuNearest(NuGeom g, int32_t* __restrict__ rows, double* __restrict__ vec) {
  const int s = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (s >= 2 * g.n) return;
  const int ii = g.xMin + (s < g.n ? s : s - g.n);
  double best = kNuFar, bx = 0., by = 0.;
  int row = -1;
  for (int k = g.slotStart[s]; k < g.slotStart[s + 1]; k++) {
    const int l = g.slotLinks[k];
    const double P0x = g.C[2 * g.BL[2 * l]], P0y = g.C[2 * g.BL[2 * l] + 1];
    const double P1x = g.C[2 * g.BL[2 * l + 1]], P1y = g.C[2 * g.BL[2 * l + 1] + 1];
    // convert_int2_rtz(fmin) - 1 and convert_int2_rtp(fmax) + 1, clamped to the lattice;
    // rows 0 and nY - 1 hold no neighbour on one side and are never taken
    const int j0 = max(int(fmin(P0y, P1y)) - 1, 1);
    const int j1 = min(int(ceil(fmax(P0y, P1y))) + 1, g.nY - 2);
    for (int jj = j0; jj <= j1; jj++) {
      if (g.nType[ii + int64_t(g.XF) * jj] & 0x0001) continue;  // M_SOLID_NODE
      double d, cx, cy;
      if (nuFoot(double(ii), double(jj), P0x, P0y, P1x, P1y, d, cx, cy) && d < best) {
        best = d;
        bx = double(ii) - cx;
        by = double(jj) - cy;
        row = jj;
      }
    }
  }
  int type = 0;
  if (row != -1) {
    if (bx < 0.) type |= 1;
    if (by < 0.) type |= 2;
  }
  rows[s] = int32_t(uint32_t(row) << 2) | type;
  vec[2 * s] = bx;
  vec[2 * s + 1] = by;
}

// getTandDT (tfdKernels.cl:878-917): the gradient towards the wall along
// x and y through an air gap dx and a soot layer dx0 - dx in series, its
// component along |vec| / |vec|, and the temperature extrapolated to the
// wall, T - dT |vec|.
__device__ inline double nuWall(const double* __restrict__ T, const double* __restrict__ dX,
                                const double* __restrict__ dX0, int64_t lin, int type, double vx,
                                double vy, int XF, int64_t D, double kSoot, double kAir,
                                double& dT) {
  const double Tn = T[lin];
  int64_t nb = (type & 1) ? 1 : -1;
  int64_t dir = ((type & 1) ? 0 : D) + lin;
  double Td = Tn - T[lin + nb];
  double dx = dX[dir], dx0 = dX0[dir];
  const double gx = kSoot * Td / (kAir * (dx0 - dx) + kSoot * dx);
  nb = (type & 2) ? XF : -XF;
  dir = ((type & 2) ? 2 * D : 3 * D) + lin;
  Td = Tn - T[lin + nb];
  dx = dX[dir];
  dx0 = dX0[dir];
  const double gy = kSoot * Td / (kAir * (dx0 - dx) + kSoot * dx);
  const double ax = fabs(vx), ay = fabs(vy);
  const double al = sqrt(ax * ax + ay * ay);
  dT = gx * (ax / al) + gy * (ay / al);
  return Tn - dT * sqrt(vx * vx + vy * vy);
}

// FD_calc_Nu: per column xMin + ix, out[5 ix ..] = bulk temperature, then
// (gradient, wall temperature) of the bottom and of the top wall, -1000
// where the wall has no node (tfdKernels.cl:919-982).  The bulk
// temperature is getTbm (:858-876): sum |u| T / sum |u| over j < height at ival + j * nY -- the
// reference steps by CHANNEL_HEIGHT (nY), not by the row length, and that
// index is kept.
template <class R>
__global__ void nuColumns(const double* __restrict__ T, const R* __restrict__ ux,
                          const R* __restrict__ uy, const double* __restrict__ dX,
                          const double* __restrict__ dX0, const int32_t* __restrict__ rows,
                          const double* __restrict__ vec, int n, int xMin, int nY, int XF,
                          int64_t D, int height, double kSoot, double kAir,
                          double* __restrict__ out) {
  const int ix = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (ix >= n) return;
  const int ii = xMin + ix;
  double Um = 0., Tm = 0.;
  for (int j = 0; j < height; j++) {
    const int64_t l = ii + int64_t(j) * nY;
    const double a = double(ux[l]), b = double(uy[l]);
    const double U = sqrt(a * a + b * b);
    Um += U;
    Tm += U * T[l];
  }
  out[5 * ix] = Tm / Um;
  for (int w = 0; w < 2; w++) {
    const int s = ix + w * n;
    const int32_t p = rows[s];
    const int jj = p >> 2;
    double dT = 0., Ts = 0.;
    if (jj != -1)
      Ts = nuWall(T, dX, dX0, ii + int64_t(XF) * jj, p & 3, vec[2 * s], vec[2 * s + 1], XF, D,
                     kSoot, kAir, dT);
    out[5 * ix + 1 + 2 * w] = jj > -1 ? dT : kNuNone;
    out[5 * ix + 2 + 2 * w] = jj > -1 ? Ts : kNuNone;
  }
}

// getNuMean (the OpenCL 1.2 form, tfdKernels.cl:984-1033): one block per
// location xloc = wavyStart + g * period; thread t takes rows 2t and 2t + 1
// and the block sums |u| T and |u| with the reference's halving tree over
// its blockDim.x (= the reference's work-group, nY / 2 rounded up to 32;
// when that is not a power of two the tree leaves some partial sums out, as
// the reference's does).  out[2 g] = sum |u| T, out[2 g + 1] = sum |u|.
template <class R>
__global__ void nuMeans(const double* __restrict__ T, const R* __restrict__ ux,
                        const R* __restrict__ uy, int wavyStart, int period, int nY, int XF,
                        double* __restrict__ out) {
  __shared__ double ut[kNuMeanMaxThreads], us[kNuMeanMaxThreads];
  const int t = int(threadIdx.x);
  const int g = int(blockIdx.x);
  const int xloc = wavyStart + g * period;
  const int st = 2 * t;
  double U1 = 0., U2 = 0., T1 = 0., T2 = 0.;
  if (st < nY) {
    const int64_t l = xloc + int64_t(st) * XF;
    const double a = double(ux[l]), b = double(uy[l]);
    U1 = sqrt(a * a + b * b);
    T1 = T[l];
  }
  if (st + 1 < nY) {
    const int64_t l = xloc + int64_t(st + 1) * XF;
    const double a = double(ux[l]), b = double(uy[l]);
    U2 = sqrt(a * a + b * b);
    T2 = T[l];
  }
  ut[t] = U1 * T1 + U2 * T2;
  us[t] = U1 + U2;
  __syncthreads();
  for (int s = int(blockDim.x) >> 1; s > 0; s >>= 1) {
    if (t < s) {
      ut[t] += ut[t + s];
      us[t] += us[t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[2 * g] = ut[0];
    out[2 * g + 1] = us[0];
  }
}

}  // namespace dev
}  // namespace kofm
