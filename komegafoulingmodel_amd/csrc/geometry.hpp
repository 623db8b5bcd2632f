// geometry.hpp -- host-side channel geometry: node types, wall distances,
// interpolated-bounce-back links, wall nodes, wall distance field and the
// wall-shear coefficients.  Built once per geometry and uploaded.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace kofm {

// nType bit flags, identical to the reference (ConstDef.h:450-480).
enum : uint16_t {
  kMSolid = 0x0001,
  kMFluid = 0x0002,
  kM0Solid = 0x0004,
  kM0Fluid = 0x0008,
  kSolidBoundary = 0x0010,
  kSolidBoundary0 = 0x0020,
  kNeswBoundary0 = 0x0040,
  kBoundary0 = 0x0080,
  kEBound = 0x0100,
  kWBound = 0x0200,
  kNBound = 0x0400,
  kSBound = 0x0800,
  kNEBound = 0x1000,
  kSWBound = 0x2000,
  kSEBound = 0x4000,
  kNWBound = 0x8000,
  kNeswBoundary = 0x0F00,
  kAnyBound = 0xFF00,
};

// D2Q9 lattice in the reference order C,E,W,N,S,NE,SW,SE,NW.
struct Lattice {
  static constexpr int cx[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
  static constexpr int cy[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
  static constexpr int rev[9] = {0, 2, 1, 4, 3, 6, 5, 8, 7};
  static constexpr uint16_t bound[9] = {0,       kEBound,  kWBound,  kNBound, kSBound,
                                        kNEBound, kSWBound, kSEBound, kNWBound};
  static constexpr double w[9] = {4. / 9.,  1. / 9.,  1. / 9.,  1. / 9., 1. / 9.,
                                  1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

struct GeometryInput {
  int nX = 0, nY = 0;
  std::vector<double> C0;  // 2*nN, x y interleaved
  std::vector<int32_t> BL; // 2*nBL
  double pipeRadius = 0.;
  double eps = 1e-5;
  double minDx = 0.001;
  int wallDSearchRad = 50;
  double xRelease = 90., xMin = 88.75, xStop = 1900.;
  double cutoff = 5.;
  int indexRadiusSearch = 30;
};

struct Geometry {
  int nX = 0, nY = 0, XF = 0;
  int64_t D = 0;
  std::vector<int16_t> nType;     // D
  std::vector<double> dX;         // 4*D  (E, W, N, S)
  std::vector<int32_t> ibb;       // nIbb, node + D*dir
  std::vector<double> ibbQ;       // nIbb
  std::vector<int32_t> ss;        // nSS
  std::vector<int32_t> ssMap;     // D
  std::vector<int32_t> ssInds;    // nX
  std::vector<uint16_t> lsMap;    // 2*XF
  std::vector<double> wallD;      // D
  std::vector<double> cfss;       // 16*nSS
  std::vector<int32_t> sInd;      // 4*nBLshear
  std::vector<double> weights;    // 4*nBLshear
  int32_t bounds[4] = {0, 0, 0, 0};
  double masses[2] = {0., 0.};
  int ssMinX = 0, ssMaxX = 0;

  // First-order sensitivity of the continuous geometry data to the y
  // coordinates of the wall points (the x coordinates stay fixed), for the
  // adjoint.  The node types, the link set and the four-nearest choices are
  // piecewise constant in the points and are not differentiated.
  std::vector<double> C0;         // 2*nN, the wall points the geometry was built from
  std::vector<int32_t> BL;        // 2*nBL
  double cutoff = 0.;             // Cutoff Radius of the shear weights
  std::vector<int32_t> ibbSeg;    // nIbb: the link cuts the wall segment (s, s+1)
  std::vector<double> ibbDqDy;    // 2*nIbb: dq/dy of points s and s+1
  std::vector<int32_t> shearLink; // nBLshear: boundary link of each shear entry
  std::vector<int32_t> wallDLink; // D: link the wall distance was taken from (-1: none)
  std::vector<int8_t> wallDCase;  // D: 0 perpendicular foot, 1 first point, 2 second point
  std::vector<int64_t> dxSlot;    // wall spacings set by a crossing: slot node + D * k
  std::vector<int32_t> dxLink;    // and the link crossed
  GeometryInput input;            // what the geometry was built from

  int nBLshear() const { return bounds[1] + bounds[3]; }
};

// Builds everything from the wall polyline.  Returns 0 or a negative
// KOFM_ERR_* code with a message in err.
int buildGeometry(const GeometryInput& in, Geometry& out, std::string& err);

// After a wall update (node types in g updated): the wall-node list and
// the wall-shear link geometry, normals and weights for the wall points C.
int refreshWallDerived(const std::vector<double>& C, Geometry& g, std::string& err);

// Adds to dy (one entry per wall point) the y-sensitivity of the
// wall distance field and of the axis wall spacings, weighted by wdBar (D
// values, one per node) and dxBar (4*D, planes E, W, N, S):
//   dy[p] += sum_n wdBar[n] dWallD[n]/dy_p + sum_s dxBar[s] d dX[s]/dy_p.
// Each value is differentiated through the wall link (or end point) that
// produced it, as the geometry build chose it; spacings held at the minimum
// spacing, and nodes farther than any link, contribute nothing.
void wallFieldGradient(const Geometry& g, const double* wdBar, const double* dxBar,
                       std::vector<double>& dy);

}  // namespace kofm
