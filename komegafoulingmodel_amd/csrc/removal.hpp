// removal.hpp -- re-entrainment of deposited particles by the wall shear
// (TR_shear_removal, trKernelsShear.cl:13-121): the particle tracer's
// consumer of the boundary-link shear.  The particle arrays and the deposit
// counters belong to the tracer and are passed in as device pointers; the
// link arrays (first point, unit normal, length, surface type) are the
// tracer's BL arrays as updateBoundaryNodes (lsKernels.cl:246-258) computes
// them from the walls, and the shear is the link shear of TR_shear_2.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace kofm {
namespace dev {

constexpr int kMaxParSizes = 32;

// particleProperties' constant arrays (particleProperties.cpp:264-330) and
// the tracer's defines (clVariablesTR.cpp:1900-1942)
struct RemovalParams {
  double tauCrit[2 * kMaxParSizes];  // paramTauCrit: per size, clean then fouled surface
  double dcoeff[kMaxParSizes];       // paramDcoeff
  double lcoeff[kMaxParSizes];       // paramLcoeff
  double mp[kMaxParSizes];           // paramMp
  int nSizes;                        // NUM_PAR_SIZES
  double dtWall;                     // DTTR_WALL
  double xMinVal, xMaxVal;           // X_MIN_VAL = X Release Pos - 0.5, X_MAX_VAL = X Stop Pos
  int xIndStart;                     // TR_X_IND_START
  int locStride;                     // FULLSIZEX_TR_PADDED
  int timerStart;                    // PAR_TIMER_START
  int nBL;
};

// Per tracer link: (shear, surface type) for the call, and the start
// point and unit normal of a re-entrained particle, (cx0, cy0, nx, ny) with
// (cx0, cy0) = first point + (ny, -nx) len / 2 + (nx, ny) 0.001, the link's
// midpoint moved 0.001 into the flow (computed on the host with the same
// operations when the walls change).
struct RemovalLinks {
  double2* ts;          // nBL: written by removalLinkShear each call
  const double* geo;    // 4 * nBL
};

struct Particles {
  double* pos;             // 2 per particle (P_pos, double2: 16-byte aligned)
  const uint32_t* numRep;  // P_Num_rep
  const int16_t* type;     // P_type (size index)
  int32_t* depFlag;        // P_Dep_Flag: link index while temporarily deposited
  uint16_t* depTimer;      // P_Dep_timer
  int32_t* loc;            // P_loc
  int32_t* timer;          // P_timer
};

// (shear, surface type) of every tracer link: the link shear output entry
// tauAt[t], 0 where TR_shear_2 writes none.
template <class R>
__global__ __launch_bounds__(256) void removalLinkShear(const R* __restrict__ tau,
                                                        const int32_t* __restrict__ tauAt,
                                                        const int16_t* __restrict__ surf, int nBL,
                                                        double2* __restrict__ ts) {
  const int t = int(blockIdx.x * blockDim.x + threadIdx.x);
  if (t >= nBL) return;
  const int at = tauAt[t];
  ts[t] = make_double2(at >= 0 ? double(tau[at]) : 0., double(surf[t]));
}

// Particles per thread: the loads of all of them are issued before any is
// used, so that a wave keeps several memory round trips in flight.
constexpr int kRemovalPer = 8;

// Particle k of the call is offset + k, k < count; a block of 256 threads
// takes 256 * kRemovalPer of them, thread t the ones t + 256 j (so that the
// lanes of a wave load and store neighbouring particles).  A particle
// whose link index or size index lies outside the arrays is left as it is
// (the reference would read past them).  The per-size constants are staged
// in LDS; a particle reads one 16-byte link record, and a second of 32 bytes
// when it re-enters the flow.
__global__ __launch_bounds__(256) void shearRemoval(RemovalParams p, RemovalLinks L, Particles P,
                                                    uint32_t* __restrict__ blDep, int64_t offset,
                                                    int64_t count) {
  __shared__ double crit[2 * kMaxParSizes], dco[kMaxParSizes], lco[kMaxParSizes],
      mass[kMaxParSizes];
  const int tid = int(threadIdx.x);
  const int64_t base = int64_t(blockIdx.x) * (256 * kRemovalPer) + tid;
  int bl[kRemovalPer], sz[kRemovalPer];
  uint16_t tm[kRemovalPer];
#pragma unroll
  for (int j = 0; j < kRemovalPer; j++) {
    const int64_t k = base + 256 * j;
    bl[j] = -1;
    sz[j] = 0;
    tm[j] = 0;
    if (k < count) {
      bl[j] = P.depFlag[k + offset];
      sz[j] = P.type[k + offset];
      tm[j] = P.depTimer[k + offset];
    }
  }
  double2 ts[kRemovalPer];
#pragma unroll
  for (int j = 0; j < kRemovalPer; j++) {
    // not deposited on a link (flag < 0), or indices outside the arrays
    if (bl[j] < 0 || bl[j] >= p.nBL || sz[j] < 0 || sz[j] >= p.nSizes) bl[j] = -1;
    ts[j] = make_double2(0., 0.);
    if (bl[j] >= 0 && uint16_t(tm[j] - 1) != 0) ts[j] = L.ts[bl[j]];
  }
  // the constants go to LDS while the particle loads and gathers above are
  // in flight
  if (tid < 2 * kMaxParSizes) crit[tid] = p.tauCrit[tid];
  if (tid < kMaxParSizes) {
    dco[tid] = p.dcoeff[tid];
    lco[tid] = p.lcoeff[tid];
    mass[tid] = p.mp[tid];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kRemovalPer; j++) {
    if (bl[j] < 0) continue;
    const int64_t i = base + 256 * j + offset;
    const int b = bl[j], s = sz[j];
    const uint16_t depTimer = uint16_t(tm[j] - 1);
    // the deposit timer ran out: the particle stays for good
    if (depTimer == 0) {
      // the reference's index, link + NUM_PAR_SIZES * size
      atomicAdd(&blDep[b + p.nSizes * s], P.numRep[i]);
      P.depFlag[i] = -2;
      P.loc[i] = -2;
      continue;
    }
    const double shear = ts[j].x;
    if (fabs(shear) <= crit[s * 2 + int(ts[j].y)]) {
      P.depTimer[i] = depTimer;
      continue;
    }
    // re-entrained: drag along the link, lift along its normal
    const double4 g = reinterpret_cast<const double4*>(L.geo)[b];
    const double nx = g.z, ny = g.w;
    const double fd = dco[s];
    const double fl = lco[s] * sqrt(shear);
    const double dx = fd * ny, dy = fd * -nx;
    const double rx = nx * fl + dx, ry = ny * fl + dy;
    const double m = mass[s];
    const double cx = g.x + rx * shear / m * p.dtWall * p.dtWall / 2.;
    const double cy = g.y + ry * shear / m * p.dtWall * p.dtWall / 2.;
    // convert_int2 (toward zero); a negative shear beyond the critical value
    // gives sqrt of it and a NaN position, whose index is taken as 0
    const int ix = cx == cx ? int(cx) : 0, iy = cy == cy ? int(cy) : 0;
    int loc = (ix - p.xIndStart) + iy * p.locStride;
    // outside the tracer's range the location is -2; the reference's flag
    // for that case is overwritten by the -1 below, and so it is here
    if (cx < p.xMinVal || cx >= p.xMaxVal) loc = -2;
    P.depFlag[i] = -1;
    P.loc[i] = loc;
    reinterpret_cast<double2*>(P.pos)[i] = make_double2(cx, cy);
    P.timer[i] = p.timerStart;
  }
}

}  // namespace dev
}  // namespace kofm
