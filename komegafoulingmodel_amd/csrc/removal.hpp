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

// the tracer's link arrays, indexed by the tracer's link index
struct RemovalLinks {
  const double* p0;      // 2 * nBL: lsc[P01ind.x]
  const double* vN;      // 2 * nBL: BL_vNvec
  const double* len;     // nBL: BL_blLen
  const int16_t* surf;   // nBL: BL_Surf_Type (int_type)
  const int32_t* tauAt;  // nBL: entry of the link shear output, -1 where none is computed
};

struct Particles {
  double* pos;             // 2 per particle
  const uint32_t* numRep;  // P_Num_rep
  const int16_t* type;     // P_type (size index)
  int32_t* depFlag;        // P_Dep_Flag: link index while temporarily deposited
  uint16_t* depTimer;      // P_Dep_timer
  int32_t* loc;            // P_loc
  int32_t* timer;          // P_timer
};

// One thread per particle offset + i, i < count.  A particle whose link
// index or size index lies outside the arrays is left as it is (the
// reference would read past them).
template <class R>
__global__ __launch_bounds__(256) void shearRemoval(RemovalParams p, RemovalLinks L,
                                                    const R* __restrict__ tau, Particles P,
                                                    uint32_t* __restrict__ blDep, int64_t offset,
                                                    int64_t count) {
  const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int64_t i = k + offset;
  const int bl = P.depFlag[i];
  if (bl < 0) return;  // not deposited on a link
  const int sz = P.type[i];
  if (bl >= p.nBL || sz < 0 || sz >= p.nSizes) return;
  const uint16_t depTimer = uint16_t(P.depTimer[i] - 1);
  // the deposit timer ran out: the particle stays for good
  if (depTimer == 0) {
    // the reference's index, link + NUM_PAR_SIZES * size
    atomicAdd(&blDep[bl + p.nSizes * sz], P.numRep[i]);
    P.depFlag[i] = -2;
    P.loc[i] = -2;
    return;
  }
  const int surf = sz * 2 + L.surf[bl];
  const int at = L.tauAt[bl];
  const double shear = at >= 0 ? double(tau[at]) : 0.;
  if (fabs(shear) <= p.tauCrit[surf]) {
    P.depTimer[i] = depTimer;
    return;
  }
  // re-entrained: drag along the link, lift along its normal, from the
  // link's midpoint moved 0.001 into the flow
  const double fd = p.dcoeff[sz];
  const double fl = p.lcoeff[sz] * sqrt(shear);
  const double nx = L.vN[2 * bl], ny = L.vN[2 * bl + 1];
  const double dx = fd * ny, dy = fd * -nx;
  const double h = L.len[bl] * 0.5;
  double cx = L.p0[2 * bl] + ny * h + nx * 0.001;
  double cy = L.p0[2 * bl + 1] + -nx * h + ny * 0.001;
  const double rx = nx * fl + dx, ry = ny * fl + dy;
  cx += rx * shear / p.mp[sz] * p.dtWall * p.dtWall / 2.;
  cy += ry * shear / p.mp[sz] * p.dtWall * p.dtWall / 2.;
  // convert_int2 (toward zero); a negative shear beyond the critical value
  // gives sqrt of it and a NaN position, whose index is taken as 0
  const int ix = cx == cx ? int(cx) : 0, iy = cy == cy ? int(cy) : 0;
  int loc = (ix - p.xIndStart) + iy * p.locStride;
  // outside the tracer's range the location is -2; the reference's flag for
  // that case is overwritten by the -1 below, and so it is here
  if (cx < p.xMinVal || cx >= p.xMaxVal) loc = -2;
  P.depFlag[i] = -1;
  P.loc[i] = loc;
  P.pos[2 * i] = cx;
  P.pos[2 * i + 1] = cy;
  P.timer[i] = p.timerStart;
}

}  // namespace dev
}  // namespace kofm
