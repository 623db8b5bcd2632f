// geometry.cpp -- host geometry builder.  Follows the reference's
// clVariablesLS::ini (clVariablesLS.cpp:569-620) step by step so that the
// node-type map, link list and wall distances are bit-identical to the
// reference algorithm; the IBB link search (a device kernel in the
// reference, lsKernels.cl:429-505) runs here once, deterministically.
#include "geometry.hpp"

#include <algorithm>
#include <cmath>
#include <unordered_map>

#include "../../include/kofm.h"

namespace kofm {
namespace {

struct V2 {
  double x, y;
};
inline V2 sub(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }
inline double len(V2 a) { return std::sqrt(a.x * a.x + a.y * a.y); }
inline int wrap(int a, int n) { return ((a % n) + n) % n; }

class Builder {
 public:
  Builder(const GeometryInput& in, Geometry& g) : in_(in), g_(g) {}

  int run(std::string& err) {
    g_.nX = in_.nX;
    g_.nY = in_.nY;
    g_.C0 = in_.C0;
    g_.BL = in_.BL;
    g_.input = in_;
    g_.cutoff = in_.cutoff;
    g_.XF = ((in_.nX + 127) / 128) * 128;
    g_.D = int64_t(g_.XF) * g_.nY;
    nN_ = int(in_.C0.size() / 2);
    nBL_ = int(in_.BL.size() / 2);
    for (int i = 0; i < nBL_ * 2; i++)
      if (in_.BL[i] >= nN_) {
        err = "boundary link index out of range of the wall points";
        return KOFM_ERR_INITIALIZING_VLS;
      }
    if (!markSolid(kM0Solid, err)) return KOFM_ERR_INITIALIZING_VLS;
    if (!boundaryFlags(true, err)) return KOFM_ERR_INITIALIZING_VLS;
    promoteFluid();
    countSolid();
    wallSpacing();
    g_.dxSlot.clear();
    g_.dxLink.clear();
    {
      std::vector<std::pair<int64_t, int32_t>> v(dxSrc_.begin(), dxSrc_.end());
      std::sort(v.begin(), v.end());
      for (const auto& e : v) {
        g_.dxSlot.push_back(e.first);
        g_.dxLink.push_back(e.second);
      }
    }
    g_.ssMinX = std::max(int(std::floor(in_.xRelease)) - 10, 0);
    g_.ssMaxX = int(std::ceil(in_.xStop)) + 10;
    if (g_.ssMaxX > g_.nX - 1) {
      err = "X Stop Pos + 10 must stay inside the lattice (ssArrInds is indexed there)";
      return KOFM_ERR_INITIALIZING_VLS;
    }
    wallNodes();
    if (!boundaryFlags(false, err)) return KOFM_ERR_INITIALIZING_VLS;
    lsMap();
    wallDistance();
    ibbLinks();
    if (!shearSetup(err)) return KOFM_ERR_INITIALIZING_VLS;
    return 0;
  }

  // The wall-node list and the shear tables again, for updated node types.
  int refresh(std::string& err) {
    nN_ = int(in_.C0.size() / 2);
    nBL_ = int(in_.BL.size() / 2);
    wallNodes();
    if (!shearSetup(err)) return KOFM_ERR_INITIALIZING_VLS;
    return 0;
  }

 private:
  const GeometryInput& in_;
  Geometry& g_;
  int nN_ = 0, nBL_ = 0;

 private:
  std::unordered_map<int64_t, int32_t> dxSrc_;  // slot -> link of the crossing that set it

  V2 pt(int idx) const { return {in_.C0[2 * idx], in_.C0[2 * idx + 1]}; }
  V2 linkP0(int bl) const { return pt(in_.BL[2 * bl]); }
  V2 linkP1(int bl) const { return pt(in_.BL[2 * bl + 1]); }
  int16_t& type(int i, int j) { return g_.nType[i + int64_t(g_.XF) * j]; }

  // iniFillMap0 (clVariablesLS.cpp:810-891).  A solid node is reset to the
  // bare solid flag (RESET_NODE_TO_M0_SOLID clears every other bit).
  bool markSolid(uint16_t solid, std::string& err) {
    g_.nType.assign(g_.D, kM0Fluid);
    for (int n = 0; n < nBL_; n++) {
      V2 c0 = linkP0(n), c1 = linkP1(n);
      V2 nrm = {c0.y - c1.y, c1.x - c0.x};
      double l = len(nrm);
      nrm = {nrm.x / l, nrm.y / l};
      int lox = int(std::floor(std::min(c0.x, c1.x))), loy = int(std::floor(std::min(c0.y, c1.y)));
      int hix = int(std::ceil(std::max(c0.x, c1.x))), hiy = int(std::ceil(std::max(c0.y, c1.y)));
      double maxx = c0.x > c1.x ? c0.x : c1.x, minx = c0.x < c1.x ? c0.x : c1.x;
      if (hix < 0 || lox >= g_.nX) continue;
      lox = std::max(lox, 0);
      hix = std::min(hix, g_.nX - 1);
      loy = std::max(loy, 0);
      hiy = std::min(hiy, g_.nY - 1);
      for (int i = lox; i <= hix; i++) {
        if (i < minx || i > maxx) continue;
        for (int j = loy; j <= hiy; j++) {
          V2 d = {double(i) - c0.x, double(j) - c0.y};
          if (dot(d, nrm) < 0) type(i, j) = int16_t(solid);
        }
      }
    }
    // column fill: solid up to the first wall node, then (after skipping 5
    // rows) solid above the next solid node found.
    for (int i = 0; i < g_.nX; i++) {
      int j = 0;
      for (;; j++) {
        if (j >= g_.nY) {
          err = "a lattice column has no wall crossing";
          return false;
        }
        if (type(i, j) & solid) break;
        type(i, j) = int16_t(solid);
      }
      j += 6;
      while (j < g_.nY) {
        bool hit = type(i, j) & solid;
        j++;
        if (hit) break;
      }
      for (; j < g_.nY; j++) type(i, j) = int16_t(solid);
    }
    return true;
  }

  // iniNodeBoundaryInfo0 / iniNodeBoundaryInfo (clVariablesLS.cpp:920-999).
  bool boundaryFlags(bool initial, std::string& err) {
    const uint16_t solid = initial ? kM0Solid : kMSolid;
    for (int i = 0; i < g_.nX; i++)
      for (int j = 0; j < g_.nY; j++) {
        if (type(i, j) & solid) continue;
        for (int m = 1; m < 9; m++) {
          int is = wrap(i + Lattice::cx[m], g_.nX), js = j + Lattice::cy[m];
          if (js < 0 || js >= g_.nY) {
            err = "Fluid node located at either y = 0 or y = nY-1. These two rows must only "
                  "contain solid nodes";
            return false;
          }
          if (!(type(is, js) & solid)) continue;
          if (initial) {
            type(i, j) |= kBoundary0;
            if (m < 5) {
              type(is, js) |= kSolidBoundary0;
              type(i, j) |= kNeswBoundary0;
            }
          } else {
            type(i, j) |= Lattice::bound[m];
            if (m < 5) type(is, js) |= kSolidBoundary;
          }
        }
      }
    return true;
  }

  // iniFillMap on a fresh run (clVariablesLS.cpp:686-708).
  void promoteFluid() {
    for (int i = 0; i < g_.nX; i++)
      for (int j = 0; j < g_.nY; j++) type(i, j) |= (type(i, j) & kM0Fluid) ? kMFluid : kMSolid;
  }

  // iniCountSolid (clVariablesLS.cpp:623-647).
  void countSolid() {
    int nSf = 0;
    for (int i = 0; i < g_.nX; i++)
      for (int j = 0; j < g_.nY; j++)
        if ((type(i, j) & kMSolid) && (type(i, j) & kM0Fluid)) nSf++;
    g_.masses[0] = double(2 * int(in_.pipeRadius) * g_.nX - nSf);
    g_.masses[1] = double(nSf);
  }

  // bcFindIntersection (clVariablesLS.cpp:144-195) with
  // bcFindIntersectionLinePlane (:197-209) and testInside (:347-356).
  bool crossing(V2 L0, V2 Ld, V2 P0, V2 P1, int cut0[2], int cut1[2], double& dist) {
    V2 p10 = sub(P1, P0);
    V2 vn = {-p10.y, p10.x};
    double den = dot(vn, Ld);
    if (den == 0.) return false;
    double distF = dot(vn, sub(P0, L0)) / den;
    V2 cut = {L0.x + Ld.x * distF, L0.y + Ld.y * distF};
    if (!(std::fabs(len(p10) - len(sub(cut, P0)) - len(sub(cut, P1))) < in_.eps)) return false;
    int ncut = int(std::ceil(distF));
    dist = 1. - (double(ncut) - distF);
    cut1[0] = int(L0.x + Ld.x * ncut);
    cut1[1] = int(L0.y + Ld.y * ncut);
    cut0[0] = cut1[0] - int(Ld.x);
    cut0[1] = cut1[1] - int(Ld.y);
    if (wrap(cut0[0], g_.nX) != cut0[0] || wrap(cut0[1], g_.nY) != cut0[1]) return false;
    // the reference's "retry one node back" branch can never succeed
    return !(type(cut0[0], cut0[1]) & kMSolid);
  }

  // inidXArrays with bcFindNodes/bcFindDirection/bcSetdXArr0
  // (clVariablesLS.cpp:650-682, 92-141, 211-286): axis directions only.
  void wallSpacing() {
    const int64_t D = g_.D;
    g_.dX.assign(4 * D, 0.);
    for (int j = 0; j < g_.nY; j++)
      for (int i = 0; i < g_.nX; i++)
        if (type(i, j) & kMFluid)
          for (int k = 0; k < 4; k++) g_.dX[i + int64_t(g_.XF) * j + D * k] = 1.;
    for (int bl = 0; bl < nBL_; bl++) {
      if (in_.BL[2 * bl] < 0) continue;
      V2 c0 = linkP0(bl), c1 = linkP1(bl);
      V2 cn = {c0.y - c1.y, c1.x - c0.x};
      int lo[2] = {int(std::floor(std::min(c0.x, c1.x))), int(std::floor(std::min(c0.y, c1.y)))};
      int hi[2] = {int(std::ceil(std::max(c0.x, c1.x))), int(std::ceil(std::max(c0.y, c1.y)))};
      if (lo[0] < 0) lo[0] = 0;
      if (lo[0] > g_.nX - 1 || hi[0] < 0) continue;
      if (hi[0] > g_.nX - 1) hi[0] = g_.nX - 1;
      if (lo[1] < 0) lo[1] = 0;
      if (lo[1] > g_.nY - 1 || hi[1] < 0) continue;
      if (hi[1] > g_.nY - 1) hi[1] = g_.nY - 1;
      for (int dir : {1, 3}) {
        V2 cd = {double(Lattice::cx[dir]), double(Lattice::cy[dir])};
        double s = dot(cn, cd);
        if (s == 0.) continue;
        if (s > 0.) {
          dir = Lattice::rev[dir];
          cd = {double(Lattice::cx[dir]), double(Lattice::cy[dir])};
        }
        // rays start on the bbox face the direction enters through
        if (Lattice::cx[dir] != 0) {
          int i = Lattice::cx[dir] == 1 ? lo[0] : hi[0];
          for (int j = lo[1]; j <= hi[1]; j++) castRay({double(i), double(j)}, cd, dir, c0, c1, bl);
        }
        if (Lattice::cy[dir] != 0) {
          int j = Lattice::cy[dir] == 1 ? lo[1] : hi[1];
          for (int i = lo[0]; i <= hi[0]; i++) castRay({double(i), double(j)}, cd, dir, c0, c1, bl);
        }
      }
    }
  }

  void castRay(V2 L0, V2 cd, int dir, V2 c0, V2 c1, int bl) {
    int a[2], b[2];
    double dist;
    if (!crossing(L0, cd, c0, c1, a, b, dist)) return;
    if (a[0] >= g_.nX || a[0] < 0 || a[1] >= g_.nY || a[1] < 0) return;
    b[0] = wrap(b[0], g_.nX);
    if (b[1] >= g_.nY || b[1] < 0) return;
    if (type(a[0], a[1]) & kM0Solid) return;
    if (!(type(b[0], b[1]) & kM0Solid)) return;
    const int64_t idx = a[0] + int64_t(g_.XF) * a[1] + g_.D * (dir - 1);
    double& slot = g_.dX[idx];
    double m = std::min(dist, slot);
    // remember the crossing that sets the value (none when held at the
    // minimum spacing), for the wall-shape sensitivity
    if (!(slot < dist)) {
      if (!(m < in_.minDx))
        dxSrc_[idx] = bl;
      else
        dxSrc_.erase(idx);
    }
    slot = std::max(m, in_.minDx);
  }

  // fillShearArray (clVariablesLS.cpp:288-322).
  void wallNodes() {
    g_.ssMap.assign(g_.D, -1);
    g_.ssInds.assign(g_.nX, 0);
    g_.ss.clear();
    for (int i = 0; i < g_.nX; i++) {
      g_.ssInds[i] = int32_t(g_.ss.size());
      if (i < g_.ssMinX || i > g_.ssMaxX) continue;
      for (int j = 0; j < g_.nY; j++) {
        if (type(i, j) & kMSolid) continue;
        bool wall = false;
        for (int k = 1; k < 5 && !wall; k++) {
          int ii = i + Lattice::cx[k];
          ii = ii >= g_.nX ? ii - g_.nX : (ii < 0 ? ii + g_.nX : ii);
          wall = type(ii, j + Lattice::cy[k]) & kMSolid;
        }
        if (!wall) continue;
        g_.ssMap[i + int64_t(g_.XF) * j] = int32_t(g_.ss.size());
        g_.ss.push_back(int32_t(i + j * g_.XF));
      }
    }
  }

  // iniLSMap (clVariablesLS.cpp:893-918).
  void lsMap() {
    g_.lsMap.assign(2 * g_.XF, 0);
    const double edge = double(g_.nX) - 1.;
    for (int b = 0; b < nBL_ / 2; b++) {
      double left = linkP0(b).x, right = linkP1(b).x;
      if (!(right < 0. || left >= edge)) g_.lsMap[int(std::ceil(left))] = uint16_t(b);
      int t = b + nBL_ / 2;
      right = linkP0(t).x;
      left = linkP1(t).x;
      if (!(right < 0. || left >= edge)) g_.lsMap[int(std::ceil(left)) + g_.XF] = uint16_t(t);
    }
  }

  // kOmega::findMinDist (kOmega.cpp:136-193): distance to the nearest link
  // of the bottom range [b0,b1) and the top range [t0,t1).
  double nearestWall(int b0, int b1, int t0, int t1, V2 p, int32_t& link, int8_t& which) const {
    double best = 1000.;
    link = -1;
    which = 0;
    for (int top = 0; top < 2; top++) {
      int lo = top ? t0 : b0, hi = top ? t1 : b1;
      for (int i = lo; i < hi; i++) {
        V2 P0 = linkP0(i), P1 = linkP1(i);
        V2 t = sub(P1, P0);
        double tl = len(t);
        double d = std::fabs(t.y * p.x - t.x * p.y + P1.x * P0.y - P1.y * P0.x) / tl;
        t = {t.x / tl, t.y / tl};
        V2 foot = {p.x - (-t.y) * d, p.y - t.x * d};
        bool before = top ? foot.x > P0.x : foot.x < P0.x;
        bool after = top ? foot.x < P1.x : foot.x > P1.x;
        if (before)
          d = len(sub(p, P0));
        else if (after)
          d = len(sub(p, P1));
        if (!(best < d)) {  // the link (and which part of it) the minimum comes from
          link = i;
          which = before ? 1 : (after ? 2 : 0);
        }
        best = best < d ? best : d;
      }
    }
    return best;
  }

  // kOmega::iniWallD (kOmega.cpp:330-349).
  void wallDistance() {
    g_.wallD.assign(g_.D, 0.);
    g_.wallDLink.assign(size_t(g_.D), -1);
    g_.wallDCase.assign(size_t(g_.D), 0);
    const int R = in_.wallDSearchRad, half = nBL_ / 2;
    for (int i = 0; i < g_.nX; i++) {
      int lb = g_.lsMap[i], lt = g_.lsMap[i + g_.XF];
      int b0 = std::max(lb - R, 0), b1 = std::min(lb + R, half);
      int t0 = std::max(lt - R, half), t1 = std::min(lt + R, nBL_);
      for (int j = 0; j < g_.nY; j++) {
        if (type(i, j) & kMSolid) continue;
        const int64_t n = i + int64_t(g_.XF) * j;
        g_.wallD[n] = nearestWall(b0, b1, t0, t1, {double(i), double(j)}, g_.wallDLink[size_t(n)],
                                  g_.wallDCase[size_t(n)]);
      }
    }
  }

  // updateIBBOnly (lsKernels.cl:429-505) + bcFindIntersectionLS (:172-215),
  // one work item after another.  The direction loop continues from the
  // reversed direction when a probe is flipped, exactly as the kernel does.
  void ibbLinks() {
    static const int flip8[8] = {1, 0, 3, 2, 5, 4, 7, 6};
    g_.ibb.clear();
    g_.ibbQ.clear();
    g_.ibbSeg.clear();
    g_.ibbDqDy.clear();
    for (int s = 0; s < nN_ - 1; s++) {
      if (s == nN_ / 2 - 1) continue;
      V2 v0 = pt(s), v1 = pt(s + 1);
      V2 t = sub(v1, v0);
      double bl = len(t);
      t = {t.x / bl, t.y / bl};
      V2 vn = {-t.y, t.x};
      int lo[2] = {int(std::min(v0.x, v1.x)), int(std::min(v0.y, v1.y))};  // truncation
      int hi[2] = {int(std::ceil(std::max(v0.x, v1.x))), int(std::ceil(std::max(v0.y, v1.y)))};
      if (lo[0] < 1) lo[0] = 1;
      if (lo[0] >= g_.nX - 1 || hi[0] < 0) continue;
      if (hi[0] >= g_.nX - 1) hi[0] = g_.nX - 2;
      if (lo[1] < 0) lo[1] = 0;
      if (lo[1] >= g_.nY || hi[1] < 0) continue;
      if (hi[1] >= g_.nY) hi[1] = g_.nY - 1;
      for (int i = lo[0]; i <= hi[0]; i++)
        for (int j = lo[1]; j <= hi[1]; j++) {
          if (type(i, j) & kMSolid) continue;
          V2 L0 = {double(i), double(j)};
          for (int dir = 0; dir < 8; dir += 2) {
            V2 Ld = {double(Lattice::cx[dir + 1]), double(Lattice::cy[dir + 1])};
            double den = dot(vn, Ld);
            if (den == 0.) continue;
            if (den > 0.) {
              dir = flip8[dir];
              Ld = {-Ld.x, -Ld.y};
              den = -den;
            }
            double dist = dot(vn, sub(v0, L0)) / den;
            if (dist < 0. || dist > 1.) continue;
            V2 cut = {L0.x + Ld.x * dist, L0.y + Ld.y * dist};
            if (std::fabs(bl - len(sub(cut, v0)) - len(sub(cut, v1))) >= 1.0e-5) continue;
            g_.ibb.push_back(int32_t(i + int64_t(g_.XF) * j + g_.D * (dir + 1)));
            g_.ibbQ.push_back(dist);
            // dist = n.(v0 - L0) / n.Ld with n = (-(y1 - y0), x1 - x0)
            const double ex = v0.x - L0.x, dxs = v1.x - v0.x;
            const double M = -(v1.y - v0.y) * Ld.x + dxs * Ld.y;
            g_.ibbSeg.push_back(s);
            g_.ibbDqDy.push_back((ex + dxs - dist * Ld.x) / M);
            g_.ibbDqDy.push_back((-ex + dist * Ld.x) / M);
          }
        }
    }
  }

  // Shear link range and per-link geometry (clVariablesTR.cpp:476-658,
  // shearStress.cpp:14-25), node normals and coefficients
  // (TR_Find_Shear_Coeffs1, trKernelsShear.cl:332-378) and the four-node
  // Gaussian interpolation (TR_Find_Shear_Coeffs2, :380-496).
  bool shearSetup(std::string& err) {
    const int half = nBL_ / 2;
    const int xstart = int(std::floor(in_.xMin)) - 1, xstop = int(std::floor(in_.xStop)) + 1;
    auto midx = [&](int i) { return int((linkP0(i).x + linkP1(i).x) / 2.); };
    int lo = -1, hi = -1;
    for (int i = 0; i < half && lo < 0; i++)
      if (double(midx(i)) >= xstart) lo = i;
    for (int i = lo + 1; lo >= 0 && i < half && hi < 0; i++)
      if (midx(i) > xstop) hi = i;
    if (lo < 0 || hi < 0) {
      err = "X Release Pos / X Stop Pos do not bracket any boundary link";
      return false;
    }
    g_.bounds[0] = lo;
    g_.bounds[1] = hi - lo + 1;
    g_.bounds[2] = half + lo;
    g_.bounds[3] = hi - lo + 1;
    // per-link first point, unit normal and length, top half left-to-right
    std::vector<V2> p0(nBL_), nv(nBL_);
    std::vector<double> bl(nBL_);
    for (int k = 0; k < nBL_; k++) {
      int i = k < half ? k : (nBL_ - 1 - k) + half;
      V2 a = linkP0(i), b = linkP1(i);
      V2 t = sub(b, a);
      double l = len(t);
      bl[k] = l;
      nv[k] = {-t.y / l, t.x / l};
      p0[k] = a;
    }
    // node normals and coefficient rows
    const int nSS = int(g_.ss.size());
    g_.cfss.assign(16 * size_t(nSS), 0.);
    for (int n = 0; n < nSS; n++) {
      V2 here = {double(g_.ss[n] % g_.XF), double(g_.ss[n] / g_.XF)};
      V2 nrm = {0., 0.};
      for (int m = std::max(n - 50, 0); m < std::min(n + 50, nSS); m++) {
        V2 there = {double(g_.ss[m] % g_.XF), double(g_.ss[m] / g_.XF)};
        if (len(sub(here, there)) > in_.cutoff) continue;
        const int t = g_.nType[g_.ss[m]];
        nrm.x += (t & kEBound) ? 1. : 0.;
        nrm.x -= (t & kWBound) ? 1. : 0.;
        nrm.y += (t & kNBound) ? 1. : 0.;
        nrm.y -= (t & kSBound) ? 1. : 0.;
      }
      double l = len(nrm);
      nrm = {nrm.x / l, nrm.y / l};
      for (int a = 0; a < 8; a++) {
        double ccx = Lattice::cx[a + 1], ccy = Lattice::cy[a + 1];
        double s = ccx * nrm.x + ccy * nrm.y;
        g_.cfss[16 * size_t(n) + a] = 3. * s * (ccx - nrm.x * s);
        g_.cfss[16 * size_t(n) + 8 + a] = 3. * s * (ccy - nrm.y * s);
      }
    }
    // four nearest wall nodes to each link midpoint
    const int nb = g_.nBLshear();
    g_.shearLink.assign(size_t(nb), 0);
    g_.sInd.assign(4 * size_t(nb), -1);
    g_.weights.assign(4 * size_t(nb), 0.);
    const double kNorm = 1. / std::sqrt(2. * 3.1415926535897931);
    for (int k = 0; k < nb; k++) {
      int b = k < g_.bounds[1] ? k + g_.bounds[0] : k + g_.bounds[2] - g_.bounds[1];
      g_.shearLink[k] = b < half ? b : (nBL_ - 1 - b) + half;
      double h = bl[b] / 2.;
      V2 mid = {p0[b].x + h * nv[b].y, p0[b].y + h * (-nv[b].x)};
      int ix = int(mid.x);
      int from = g_.ssInds[std::max(g_.ssMinX, ix - in_.indexRadiusSearch)];
      int to = g_.ssInds[std::min(g_.ssMaxX, ix + in_.indexRadiusSearch + 1)];
      // keep the four smallest distances, largest first
      double dr[4] = {100., 100., 100., 100.};
      int id[4] = {-1, -1, -1, -1};
      for (int m = from; m < to; m++) {
        V2 there = {double(g_.ss[m] % g_.XF), double(g_.ss[m] / g_.XF)};
        double d = len(sub(mid, there));
        if (!(d < dr[0] && d < in_.cutoff)) continue;
        dr[0] = d;
        id[0] = m;
        for (int s = 0; s < 3 && dr[s] < dr[s + 1]; s++) {
          std::swap(dr[s], dr[s + 1]);
          std::swap(id[s], id[s + 1]);
        }
      }
      double wt[4];
      for (int s = 0; s < 4; s++) {
        double r = dr[s] / in_.cutoff;
        wt[s] = kNorm * std::exp(-0.5 * r * r);
      }
      // sort the four by wall-node index (five-comparator network)
      static const int pairs[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
      for (auto& pr : pairs)
        if (id[pr[0]] > id[pr[1]]) {
          std::swap(id[pr[0]], id[pr[1]]);
          std::swap(wt[pr[0]], wt[pr[1]]);
        }
      double sum = ((wt[0] + wt[1]) + wt[2]) + wt[3];
      for (int s = 0; s < 4; s++) {
        g_.weights[4 * size_t(k) + s] = wt[s] / sum;
        g_.sInd[4 * size_t(k) + s] = id[s];
      }
    }
    (void)err;
    return true;
  }
};

}  // namespace

int buildGeometry(const GeometryInput& in, Geometry& out, std::string& err) {
  if (in.nX < 4 || in.nY < 4 || in.C0.size() < 4 || in.BL.size() < 2) {
    err = "lattice or wall description too small";
    return KOFM_ERR_ARG;
  }
  Builder b(in, out);
  return b.run(err);
}

int refreshWallDerived(const std::vector<double>& C, Geometry& g, std::string& err) {
  GeometryInput in = g.input;
  in.C0 = C;
  Builder b(in, g);
  return b.refresh(err);
}

void wallFieldGradient(const Geometry& g, const double* wdBar, const double* dxBar,
                       std::vector<double>& dy) {
  const auto& in = g.input;
  dy.resize(in.C0.size() / 2, 0.);
  auto pt = [&](int idx) { return V2{in.C0[2 * size_t(idx)], in.C0[2 * size_t(idx) + 1]}; };
  auto add = [&](int link, double w, double d0, double d1) {
    dy[size_t(in.BL[2 * size_t(link)])] += w * d0;
    dy[size_t(in.BL[2 * size_t(link) + 1])] += w * d1;
  };
  if (wdBar)
    for (int64_t n = 0; n < g.D; n++) {
      const int link = g.wallDLink[size_t(n)];
      if (wdBar[n] == 0. || link < 0 || (g.nType[size_t(n)] & kMSolid)) continue;
      const V2 p = {double(n % g.XF), double(n / g.XF)};
      const V2 P0 = pt(in.BL[2 * size_t(link)]), P1 = pt(in.BL[2 * size_t(link) + 1]);
      const int which = g.wallDCase[size_t(n)];
      if (which == 1) {  // distance to the first point
        const double d = len(sub(p, P0));
        if (d > 0.) add(link, wdBar[n], -(p.y - P0.y) / d, 0.);
      } else if (which == 2) {
        const double d = len(sub(p, P1));
        if (d > 0.) add(link, wdBar[n], 0., -(p.y - P1.y) / d);
      } else {
        // |cr| / |P1 - P0| with cr = t.y p.x - t.x p.y + P1.x P0.y - P1.y P0.x
        const V2 t = sub(P1, P0);
        const double tl = len(t);
        const double cr = t.y * p.x - t.x * p.y + P1.x * P0.y - P1.y * P0.x;
        const double sg = cr < 0. ? -1. : 1.;
        const double c3 = std::fabs(cr) * t.y / (tl * tl * tl);
        add(link, wdBar[n], sg * (P1.x - p.x) / tl + c3, sg * (p.x - P0.x) / tl - c3);
      }
    }
  if (dxBar)
    for (size_t k = 0; k < g.dxSlot.size(); k++) {
      const int64_t slot = g.dxSlot[k];
      const int64_t node = slot % g.D;
      const int dir = int(slot / g.D) + 1;
      if (dxBar[slot] == 0. || (g.nType[size_t(node)] & kMSolid)) continue;
      // the crossing cut = a + c dist of the link (c0, c1): with
      // M = -(y1 - y0) c.x + (x1 - x0) c.y, d dist/dy0 = (x1 - cut.x) / M and
      // d dist/dy1 = (cut.x - x0) / M
      const int link = g.dxLink[k];
      const V2 c0 = pt(in.BL[2 * size_t(link)]), c1 = pt(in.BL[2 * size_t(link) + 1]);
      const double cdx = Lattice::cx[dir], cdy = Lattice::cy[dir];
      const double cutx = double(node % g.XF) + cdx * g.dX[size_t(slot)];
      const double M = -(c1.y - c0.y) * cdx + (c1.x - c0.x) * cdy;
      add(link, dxBar[slot], (c1.x - cutx) / M, (cutx - c0.x) / M);
    }
}

}  // namespace kofm
