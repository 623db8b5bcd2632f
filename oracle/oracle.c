/*
 * oracle.c -- FP64 CPU restatement of the reference lattice-Boltzmann + k-omega
 * time step (zmills04/kOmegaFoulingModel), used ONLY as test infrastructure:
 * the parity tests, the smoke check and the CPU-baseline leg of bench.py call
 * it as the checker.  Nothing in the product library links, loads or calls it.
 *
 * Every function cites the reference file:line whose algorithm it restates.
 * Paths are relative to the reference checkout; "src/" is "source/".
 *
 * Layout (as the reference, Heading.cl:413-415): SoA, index
 *   x + XF*y + XF*nY*dir,  XF = ceil(nX/128)*128, D = XF*nY.
 * Direction order C,E,W,N,S,NE,SW,SE,NW (Heading.cl:197-215).
 *
 * Deliberate differences from the reference, each pinned by a test:
 *  - the BGK/EDM collision is written in product form
 *      f <- f - w (f - feq(u)) + feq(u+du) - feq(u)
 *    instead of the reference's expanded polynomial (lbKernels.cl:224-287,
 *    464-563); the two agree to FP64 rounding (feq is the product form of
 *    clVariablesLB.cpp:77-117 in both);
 *  - the padding column x = nX is not collided (lbKernels.cl:189 collides it
 *    when nX % 128 != 0; every supported configuration has nX % 128 == 0);
 *  - interpolated bounce-back reads all its inputs before writing (the
 *    reference's parallel lbIBB, lbKernels.cl:640-668, does the same for every
 *    link whose inputs are not another link's output);
 *  - the implicit k/omega systems are solved by BiCGStab to a relative
 *    residual of 1e-12 instead of GMRES(10) to 1e-4 (GMRESSolver.cpp:227-355);
 *    the converged answer is fixed by the assembled system.
 *
 * Parity status of each part is recorded in the test that pins it
 * (tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* nType bit flags (ConstDef.h:450-480, Heading.cl:223-251). */
#define M_SOLID_NODE 0x0001
#define M_FLUID_NODE 0x0002
#define M0_SOLID_NODE 0x0004
#define M0_FLUID_NODE 0x0008
#define SOLID_BOUNDARY_NODE 0x0010
#define SOLID_BOUNDARY_NODE0 0x0020
#define NESW_BOUNDARY_NODE0 0x0040
#define BOUNDARY_NODE0 0x0080
#define E_BOUND 0x0100
#define W_BOUND 0x0200
#define N_BOUND 0x0400
#define S_BOUND 0x0800
#define NE_BOUND 0x1000
#define SW_BOUND 0x2000
#define SE_BOUND 0x4000
#define NW_BOUND 0x8000
#define NESW_BOUNDARY_NODE 0x0F00
#define BOUNDARY_NODE 0xFF00

#define OPTION_SAVE_MACRO_FIELDS 0x2

/* k-omega SST constants (Heading.cl:396-410). */
#define KO_BETA_STAR 0.09
#define KO_SIGMA_K1 0.85
#define KO_SIGMA_K2 1.0
#define KO_SIGMA_O1 0.5
#define KO_SIGMA_O2 0.856
#define KO_A1 0.31
#define KO_KAPPA 0.41
#define KO_C_MU_0_25 0.5477225575
#define KO_GAMMA1 (5. / 9.)
#define KO_GAMMA2 0.44

static const int CX[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
static const int CY[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
static const int REV[9] = {0, 2, 1, 4, 3, 6, 5, 8, 7};
static const int BOUND[9] = {0, E_BOUND, W_BOUND, N_BOUND, S_BOUND,
                             NE_BOUND, SW_BOUND, SE_BOUND, NW_BOUND};

static int imod(int a, int b) { return ((a % b) + b) % b; }

/* ------------------------------------------------------------------------ */
/* Geometry                                                                  */
/* ------------------------------------------------------------------------ */

/* Host min2/max2 (clVariablesLS.cpp:330-344): floor of the min, ceil of the max. */
static void host_bbox(const double *c0, const double *c1, int *mn, int *mx) {
  double x = c0[0], y = c0[1];
  if (c1[0] < x) x = c1[0];
  if (c1[1] < y) y = c1[1];
  mn[0] = (int)floor(x);
  mn[1] = (int)floor(y);
  x = c0[0];
  y = c0[1];
  if (c1[0] > x) x = c1[0];
  if (c1[1] > y) y = c1[1];
  mx[0] = (int)ceil(x);
  mx[1] = (int)ceil(y);
}

/* Marks nodes on the solid side of each wall link inside its bounding box,
 * then fills each column below the first and above the second solid node.
 * iniFillMap0 clVariablesLS.cpp:810-891 (flag = M0_SOLID_NODE).  The
 * reference's RESET_NODE_TO_M0_SOLID (Heading.cl:299-300) clears every other
 * bit ("&= !flag" is "&= 0"), so a solid node holds exactly M0_SOLID_NODE. */
static int fill_map0(int nX, int XF, int nY, const double *C0, const int32_t *BL,
                     int nBL, int16_t *nType) {
  const long D = (long)XF * nY;
  for (long g = 0; g < D; g++) nType[g] = M0_FLUID_NODE;
  for (int n = 0; n < nBL; n++) {
    const double *c0 = C0 + 2 * BL[2 * n], *c1 = C0 + 2 * BL[2 * n + 1];
    double cn[2] = {c0[1] - c1[1], c1[0] - c0[0]};
    double len = sqrt(cn[0] * cn[0] + cn[1] * cn[1]);
    cn[0] = cn[0] / len;
    cn[1] = cn[1] / len;
    int mn[2], mx[2];
    host_bbox(c0, c1, mn, mx);
    double maxx = c0[0] > c1[0] ? c0[0] : c1[0];
    double minx = c0[0] < c1[0] ? c0[0] : c1[0];
    if (mx[0] < 0 || mn[0] >= nX) continue;
    if (mn[0] < 0) mn[0] = 0;
    if (mx[0] >= nX) mx[0] = nX - 1;
    if (mn[1] < 0) mn[1] = 0;
    if (mx[1] >= nY) mx[1] = nY - 1;
    for (int i = mn[0]; i < mx[0] + 1; i++)
      for (int j = mn[1]; j < mx[1] + 1; j++) {
        if (i < minx || i > maxx) continue;
        double v0 = (double)i - c0[0], v1 = (double)j - c0[1];
        if (v0 * cn[0] + v1 * cn[1] < 0) nType[i + XF * j] = M0_SOLID_NODE;
      }
  }
  for (int i = 0; i < nX; i++) {
    int ind = 0, flag = 0;
    while (flag == 0) {
      if (ind >= nY) return -1; /* column without a wall */
      if (nType[i + XF * ind] & M0_SOLID_NODE)
        flag = 1;
      else
        nType[i + XF * ind] = M0_SOLID_NODE;
      ind++;
    }
    ind += 5;
    flag = 0;
    while (flag == 0 && ind < nY) {
      if (nType[i + XF * ind] & M0_SOLID_NODE) flag = 1;
      ind++;
    }
    for (int j = ind; j < nY; j++) nType[i + XF * j] = M0_SOLID_NODE;
  }
  return 0;
}

/* iniNodeBoundaryInfo0 clVariablesLS.cpp:920-951, iniFillMap (fresh run)
 * :686-708, iniNodeBoundaryInfo :953-999 (bound bits only; the dX printf
 * check has no effect).  Returns -2 when a fluid node sits on row 0 or nY-1
 * (ERROR_INITIALIZING_VLS, :935-937). */
static int boundary_info(int nX, int XF, int nY, int16_t *nType, int pass) {
  for (int i = 0; i < nX; i++)
    for (int j = 0; j < nY; j++) {
      int g = i + XF * j;
      int solid_flag = pass == 0 ? M0_SOLID_NODE : M_SOLID_NODE;
      if (nType[g] & solid_flag) continue;
      for (int m = 1; m < 9; m++) {
        int is = imod(i + CX[m], nX), js = j + CY[m];
        if (js < 0 || js >= nY) return -2;
        int gs = is + XF * js;
        if (nType[gs] & solid_flag) {
          if (pass == 0) {
            nType[g] |= BOUNDARY_NODE0;
            if (m < 5) {
              nType[gs] |= SOLID_BOUNDARY_NODE0;
              nType[g] |= NESW_BOUNDARY_NODE0;
            }
          } else {
            nType[g] |= BOUND[m];
            if (m < 5) nType[gs] |= SOLID_BOUNDARY_NODE;
          }
        }
      }
    }
  return 0;
}

/* Line/segment intersection helpers, clVariablesLS.cpp:144-209, 347-356. */
static int line_plane(double *vc, double *dist, const double *L0, const double *Ld,
                      const double *P0, const double *P1) {
  double p10[2] = {P1[0] - P0[0], P1[1] - P0[1]};
  double vn[2] = {-p10[1], p10[0]};
  double den = vn[0] * Ld[0] + vn[1] * Ld[1];
  if (den == 0.) return 0;
  double pl[2] = {P0[0] - L0[0], P0[1] - L0[1]};
  *dist = (vn[0] * pl[0] + vn[1] * pl[1]) / den;
  vc[0] = L0[0] + Ld[0] * *dist;
  vc[1] = L0[1] + Ld[1] * *dist;
  return 1;
}

static int test_inside(const double *vd, const double *v0, const double *v1, double eps) {
  double a[2] = {v1[0] - v0[0], v1[1] - v0[1]};
  double b[2] = {vd[0] - v0[0], vd[1] - v0[1]};
  double c[2] = {vd[0] - v1[0], vd[1] - v1[1]};
  double la = sqrt(a[0] * a[0] + a[1] * a[1]);
  double lb = sqrt(b[0] * b[0] + b[1] * b[1]);
  double lc = sqrt(c[0] * c[0] + c[1] * c[1]);
  return fabs(la - lb - lc) < eps;
}

/* bcFindIntersection clVariablesLS.cpp:144-195.  Returns the fluid node before
 * the wall (cut0), the node past it (cut1) and the fractional distance. */
static int find_intersection(int nX, int nY, int XF, const int16_t *nType, double eps,
                             int *cut0, int *cut1, double *dist, const double *L0,
                             const double *Ld, const double *P0, const double *P1) {
  double vc[2], distF;
  if (!line_plane(vc, &distF, L0, Ld, P0, P1)) return 0;
  if (!test_inside(vc, P0, P1, eps)) return 0;
  int n_cut = (int)ceil(distF);
  *dist = 1. - ((double)n_cut - distF);
  cut1[0] = (int)(L0[0] + Ld[0] * n_cut);
  cut1[1] = (int)(L0[1] + Ld[1] * n_cut);
  cut0[0] = cut1[0] - (int)Ld[0];
  cut0[1] = cut1[1] - (int)Ld[1];
  int ii0x = imod(cut0[0], nX), ii0y = imod(cut0[1], nY);
  if (ii0x != cut0[0] || ii0y != cut0[1]) return 0;
  /* A solid "node before the wall" never yields an entry: the reference's
   * retry with n_cut-1 compares against the old node and always fails. */
  if (nType[ii0x + XF * ii0y] & M_SOLID_NODE) return 0;
  return 1;
}

/* inidXArrays clVariablesLS.cpp:650-682 with bcFindNodes :211-246,
 * bcFindDirection :92-141 and bcSetdXArr0 :268-286 (dir = 1 and 3 only,
 * IN_KERNEL_IBB unset, ConstDef.h:169). dX layout: x + XF*y + D*k, k=E,W,N,S. */
static void dx_arrays(int nX, int XF, int nY, const double *C0, const int32_t *BL,
                      int nBL, const int16_t *nType, double eps, double mindx, double *dX) {
  const long D = (long)XF * nY;
  for (long g = 0; g < 4 * D; g++) dX[g] = 0.;
  for (int j = 0; j < nY; j++)
    for (int i = 0; i < nX; i++)
      if (nType[i + XF * j] & M_FLUID_NODE)
        for (int k = 0; k < 4; k++) dX[i + XF * j + D * k] = 1.;
  for (int bl = 0; bl < nBL; bl++) {
    int n0 = BL[2 * bl], n1 = BL[2 * bl + 1];
    if (n0 < 0) continue;
    const double *c0 = C0 + 2 * n0, *c1 = C0 + 2 * n1;
    double cn[2] = {c0[1] - c1[1], c1[0] - c0[0]};
    int mn[2], mx[2];
    host_bbox(c0, c1, mn, mx);
    if (mn[0] < 0) mn[0] = 0;
    if (mn[0] > nX - 1) continue;
    if (mx[0] < 0) continue;
    if (mx[0] > nX - 1) mx[0] = nX - 1;
    if (mn[1] < 0) mn[1] = 0;
    if (mn[1] > nY - 1) continue;
    if (mx[1] < 0) continue;
    if (mx[1] > nY - 1) mx[1] = nY - 1;
    for (int dir0 = 1; dir0 < 4; dir0 += 2) {
      int dir = dir0;
      double cd[2] = {(double)CX[dir], (double)CY[dir]};
      double cnd = cn[0] * cd[0] + cn[1] * cd[1];
      if (cnd == 0.) continue;
      if (cnd > 0.) {
        dir = REV[dir];
        cd[0] = CX[dir];
        cd[1] = CY[dir];
      }
      for (int axis = 0; axis < 2; axis++) {
        int c = axis == 0 ? CX[dir] : CY[dir];
        if (c == 0) continue;
        int fixed = axis == 0 ? (c == 1 ? mn[0] : mx[0]) : (c == 1 ? mn[1] : mx[1]);
        int lo = axis == 0 ? mn[1] : mn[0], hi = axis == 0 ? mx[1] : mx[0];
        for (int v = lo; v <= hi; v++) {
          double L0[2] = {axis == 0 ? (double)fixed : (double)v,
                          axis == 0 ? (double)v : (double)fixed};
          int cut0[2], cut1[2];
          double dist;
          if (!find_intersection(nX, nY, XF, nType, eps, cut0, cut1, &dist, L0, cd, c0, c1))
            continue;
          /* bcSetdXArr0 */
          if (cut0[0] >= nX || cut0[0] < 0 || cut0[1] >= nY || cut0[1] < 0) continue;
          cut1[0] = imod(cut1[0], nX);
          if (cut1[1] >= nY || cut1[1] < 0) continue;
          if (nType[cut0[0] + XF * cut0[1]] & M0_SOLID_NODE) continue;
          if ((nType[cut1[0] + XF * cut1[1]] & M0_SOLID_NODE) == 0) continue;
          double *d = dX + cut0[0] + XF * cut0[1] + D * (dir - 1);
          double v2 = dist < *d ? dist : *d;
          *d = v2 > mindx ? v2 : mindx;
        }
      }
    }
  }
}

/* fillShearArray clVariablesLS.cpp:288-322: fluid nodes with a solid NESW
 * neighbour inside [minX, maxX], column-major order. Returns the count. */
static int shear_nodes(int nX, int XF, int nY, const int16_t *nType, int minX, int maxX,
                       int32_t *ssArr, int32_t *ssIndMap, int32_t *ssInds) {
  const long D = (long)XF * nY;
  for (long g = 0; g < D; g++) ssIndMap[g] = -1;
  int n = 0;
  for (int i = 0; i < nX; i++) {
    ssInds[i] = n;
    for (int j = 0; j < nY; j++) {
      if (nType[i + XF * j] & M_SOLID_NODE) continue;
      int added = 0;
      for (int k = 0; k < 4; k++) {
        int ii = i + CX[k + 1];
        ii = ii >= nX ? ii - nX : (ii < 0 ? ii + nX : ii);
        int jj = j + CY[k + 1];
        if ((nType[ii + XF * jj] & M_SOLID_NODE) && i >= minX && i <= maxX && !added) {
          ssIndMap[i + XF * j] = n;
          ssArr[n++] = i + j * XF;
          added = 1;
        }
      }
    }
  }
  return n;
}

/* iniLSMap clVariablesLS.cpp:893-918 (lsMap(i,k) = base[i + XF*k]). */
static void ls_map(int nX, int XF, const double *C0, const int32_t *BL, int nBL, uint16_t *lsMap) {
  memset(lsMap, 0, sizeof(uint16_t) * 2 * XF);
  for (int ib = 0; ib < nBL / 2; ib++) {
    double cx1 = C0[2 * BL[2 * ib]], cx2 = C0[2 * BL[2 * ib + 1]];
    if (!(cx2 < 0. || cx1 >= (double)nX - 1.)) lsMap[(int)ceil(cx1)] = (uint16_t)ib;
    int it = ib + nBL / 2;
    cx2 = C0[2 * BL[2 * it]];
    cx1 = C0[2 * BL[2 * it + 1]];
    if (!(cx2 < 0. || cx1 >= (double)nX - 1.)) lsMap[(int)ceil(cx1) + XF] = (uint16_t)it;
  }
}

/* kOmega::findMinDist kOmega.cpp:136-193. */
static double min_dist(const double *C, const int32_t *BL, int b0, int b1, int t0, int t1,
                       const double *nl) {
  double dmin = 1000.;
  for (int pass = 0; pass < 2; pass++) {
    int lo = pass == 0 ? b0 : t0, hi = pass == 0 ? b1 : t1;
    for (int i = lo; i < hi; i++) {
      const double *P0 = C + 2 * BL[2 * i], *P1 = C + 2 * BL[2 * i + 1];
      double vT[2] = {P1[0] - P0[0], P1[1] - P0[1]};
      double mag = sqrt(vT[0] * vT[0] + vT[1] * vT[1]);
      double d = fabs(vT[1] * nl[0] - vT[0] * nl[1] + P1[0] * P0[1] - P1[1] * P0[0]) / mag;
      vT[0] = vT[0] / mag;
      vT[1] = vT[1] / mag;
      double vN[2] = {-vT[1], vT[0]};
      double pc[2] = {nl[0] - vN[0] * d, nl[1] - vN[1] * d};
      int before = pass == 0 ? (pc[0] < P0[0]) : (pc[0] > P0[0]);
      int after = pass == 0 ? (pc[0] > P1[0]) : (pc[0] < P1[0]);
      if (before) {
        double a = nl[0] - P0[0], b = nl[1] - P0[1];
        d = sqrt(a * a + b * b);
      } else if (after) {
        double a = nl[0] - P1[0], b = nl[1] - P1[1];
        d = sqrt(a * a + b * b);
      }
      dmin = dmin < d ? dmin : d;
    }
  }
  return dmin;
}

/* kOmega::iniWallD kOmega.cpp:330-349. */
static void wall_dist(int nX, int XF, int nY, const double *C, const int32_t *BL, int nBL,
                      const int16_t *nType, const uint16_t *lsMap, int rad, double *WallD) {
  const long D = (long)XF * nY;
  for (long g = 0; g < D; g++) WallD[g] = 0.;
  for (int i = 0; i < nX; i++) {
    int lb = lsMap[i], lt = lsMap[i + XF];
    int b0 = lb - rad > 0 ? lb - rad : 0, b1 = lb + rad < nBL / 2 ? lb + rad : nBL / 2;
    int t0 = lt - rad > nBL / 2 ? lt - rad : nBL / 2, t1 = lt + rad < nBL ? lt + rad : nBL;
    for (int j = 0; j < nY; j++) {
      if (nType[i + XF * j] & M_SOLID_NODE) continue;
      double nl[2] = {(double)i, (double)j};
      WallD[i + XF * j] = min_dist(C, BL, b0, b1, t0, t1, nl);
    }
  }
}

/* updateIBBOnly lsKernels.cl:429-505 with bcFindIntersectionLS :172-215,
 * run as one sequential work-item loop (the reference appends through an
 * atomic counter, so only the set of entries is defined).  Note the
 * direction loop increments from the possibly reversed direction. */
static int ibb_links(int nX, int XF, int nY, const double *C, int nN, const int16_t *nType,
                     int32_t *ibb, double *q, int cap) {
  static const int RD8[8] = {1, 0, 3, 2, 5, 4, 7, 6};
  const long D = (long)XF * nY;
  int count = 0;
  for (int i = 0; i < nN - 1; i++) {
    if (i == nN / 2 - 1) continue;
    const double *v0 = C + 2 * i, *v1 = C + 2 * i + 2;
    double vT[2] = {v1[0] - v0[0], v1[1] - v0[1]};
    double blLen = sqrt(vT[0] * vT[0] + vT[1] * vT[1]);
    vT[0] /= blLen;
    vT[1] /= blLen;
    double vN[2] = {-vT[1], vT[0]};
    /* device min2/max2 (Heading.cl:493-506): truncation / ceil */
    double mnx = v1[0] < v0[0] ? v1[0] : v0[0], mny = v1[1] < v0[1] ? v1[1] : v0[1];
    double mxx = v1[0] > v0[0] ? v1[0] : v0[0], mxy = v1[1] > v0[1] ? v1[1] : v0[1];
    int mn[2] = {(int)mnx, (int)mny}, mx[2] = {(int)ceil(mxx), (int)ceil(mxy)};
    if (mn[0] < 1) mn[0] = 1;
    if (mn[0] >= nX - 1) continue;
    if (mx[0] < 0) continue;
    if (mx[0] >= nX - 1) mx[0] = nX - 2;
    if (mn[1] < 0) mn[1] = 0;
    if (mn[1] >= nY) continue;
    if (mx[1] < 0) continue;
    if (mx[1] >= nY) mx[1] = nY - 1;
    for (int ii = mn[0]; ii <= mx[0]; ii++)
      for (int jj = mn[1]; jj <= mx[1]; jj++) {
        int g = ii + XF * jj;
        if (nType[g] & M_SOLID_NODE) continue;
        double L0[2] = {(double)ii, (double)jj};
        for (int dir = 0; dir < 8; dir += 2) {
          double Ld[2] = {(double)CX[dir + 1], (double)CY[dir + 1]};
          double den = vN[0] * Ld[0] + vN[1] * Ld[1];
          if (den == 0.) continue;
          if (den > 0.) {
            dir = RD8[dir];
            Ld[0] = -Ld[0];
            Ld[1] = -Ld[1];
            den = -den;
          }
          double dist = (vN[0] * (v0[0] - L0[0]) + vN[1] * (v0[1] - L0[1])) / den;
          if (dist < 0. || dist > 1.) continue;
          double cut[2] = {L0[0] + Ld[0] * dist, L0[1] + Ld[1] * dist};
          double a0 = cut[0] - v0[0], a1 = cut[1] - v0[1];
          double b0 = cut[0] - v1[0], b1 = cut[1] - v1[1];
          if (fabs(blLen - sqrt(a0 * a0 + a1 * a1) - sqrt(b0 * b0 + b1 * b1)) >= 1.0e-5) continue;
          if (count < cap) {
            ibb[count] = (int32_t)(g + D * (dir + 1));
            q[count] = dist;
          }
          count++;
        }
      }
  }
  return count;
}

/* Full host geometry build in the reference order (clVariablesLS::ini,
 * clVariablesLS.cpp:569-620; kOmega::iniWallD; clVariablesLS::iniIBBArrays
 * :1085-1164).  Returns 0, or a negative error code. */
int orc_geometry(int nX, int XF, int nY, const double *C0, int nN, const int32_t *BL,
                 int nBL, double eps, double mindx, int ssMinX, int ssMaxX, int wallDRad,
                 int16_t *nType, double *dX, int32_t *ssArr, int32_t *ssIndMap,
                 int32_t *ssInds, int *nSS, uint16_t *lsMap, double *WallD, int32_t *ibb,
                 double *ibbq, int ibbCap, int *nIbb) {
  /* ssArrInds is indexed up to ssMaxX (TR_Find_Shear_Coeffs2), so the shear
   * window has to end inside the lattice */
  if (ssMinX < 0 || ssMaxX > nX - 1) return -5;
  int rc = fill_map0(nX, XF, nY, C0, BL, nBL, nType);
  if (rc) return rc;
  rc = boundary_info(nX, XF, nY, nType, 0);
  if (rc) return rc;
  for (int i = 0; i < nX; i++)
    for (int j = 0; j < nY; j++) {
      int16_t *t = nType + i + XF * j;
      if (*t & M0_FLUID_NODE)
        *t |= M_FLUID_NODE;
      else
        *t |= M_SOLID_NODE;
    }
  dx_arrays(nX, XF, nY, C0, BL, nBL, nType, eps, mindx, dX);
  *nSS = shear_nodes(nX, XF, nY, nType, ssMinX, ssMaxX, ssArr, ssIndMap, ssInds);
  rc = boundary_info(nX, XF, nY, nType, 1);
  if (rc) return rc;
  ls_map(nX, XF, C0, BL, nBL, lsMap);
  wall_dist(nX, XF, nY, C0, BL, nBL, nType, lsMap, wallDRad, WallD);
  *nIbb = ibb_links(nX, XF, nY, C0, nN, nType, ibb, ibbq, ibbCap);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Wall-shear coefficients                                                   */
/* ------------------------------------------------------------------------ */

/* TR_Find_Shear_Coeffs1 trKernelsShear.cl:332-378: node normal from the NESW
 * bound flags of wall nodes within +-50 list entries and the cutoff radius,
 * CFSS[16*ii + a] (x) and [16*ii + 8 + a] (y) for the eight moving directions. */
void orc_shear_coeffs1(int XF, const int32_t *ssArr, int nSS, const int16_t *nType,
                       double cutoff, double *CFSS) {
  for (int ii = 0; ii < nSS; ii++) {
    int g = ssArr[ii];
    double nx0 = g % XF, ny0 = g / XF;
    double n[2] = {0., 0.};
    int lo = ii - 50 > 0 ? ii - 50 : 0, hi = ii + 50 < nSS ? ii + 50 : nSS;
    for (int j = lo; j < hi; j++) {
      int h = ssArr[j];
      double dx = nx0 - (double)(h % XF), dy = ny0 - (double)(h / XF);
      if (sqrt(dx * dx + dy * dy) > cutoff) continue;
      int t = nType[h];
      if (t & E_BOUND) n[0] += 1.;
      if (t & W_BOUND) n[0] -= 1.;
      if (t & N_BOUND) n[1] += 1.;
      if (t & S_BOUND) n[1] -= 1.;
    }
    double len = sqrt(n[0] * n[0] + n[1] * n[1]);
    n[0] /= len;
    n[1] /= len;
    for (int a = 0; a < 8; a++) {
      double cx = CX[a + 1], cy = CY[a + 1];
      double s = cx * n[0] + cy * n[1];
      CFSS[16 * ii + a] = 3. * s * (cx - n[0] * s);
      CFSS[16 * ii + 8 + a] = 3. * s * (cy - n[1] * s);
    }
  }
}

/* Shear boundary-link range and per-link geometry: clVariablesTR::iniBLandNodI
 * clVariablesTR.cpp:476-542 (range), iniBLinksBottom :544-599 and
 * iniBLinksTop :601-658 (the top half is stored left-to-right), loadParams
 * :1311-1371 (xMinVal = xRelease - 1.25 by default); shearStress.cpp:14-25
 * (sizes, inclusive ranges).  out: bounds[4] = {BL_BOT_START, NUM_BL_BOT,
 * BL_TOP_START, NUM_BL_TOP}; P0[2*nBL], tangent-ordered first point;
 * vN[2*nBL]; blLen[nBL]. */
int orc_shear_bl_setup(const double *C, const int32_t *BL, int nBL, double xRelease,
                       double xMin, double xStop, int32_t *bounds, double *P0, double *vN,
                       double *blLen) {
  int xstart = (int)floor(xMin) - 1, xstop = (int)floor(xStop) + 1;
  int minb = -1, maxb = -1;
  for (int i = 0; i < nBL / 2; i++) {
    double m = (double)(int)((C[2 * BL[2 * i]] + C[2 * BL[2 * i + 1]]) / 2.);
    if (m >= xstart) {
      minb = i;
      break;
    }
  }
  if (minb < 0) return -3;
  for (int i = minb + 1; i < nBL / 2; i++) {
    int m = (int)((C[2 * BL[2 * i]] + C[2 * BL[2 * i + 1]]) / 2.);
    if (m > xstop) {
      maxb = i;
      break;
    }
  }
  if (maxb < 0) return -3;
  bounds[0] = minb;
  bounds[1] = maxb - minb + 1;
  bounds[2] = nBL / 2 + minb;
  bounds[3] = maxb - minb + 1;
  for (int ii = 0; ii < nBL; ii++) {
    int i = ii < nBL / 2 ? ii : (nBL - 1 - ii) + nBL / 2;
    const double *a = C + 2 * BL[2 * i], *b = C + 2 * BL[2 * i + 1];
    double t[2] = {b[0] - a[0], b[1] - a[1]};
    double len = sqrt(t[0] * t[0] + t[1] * t[1]);
    blLen[ii] = len;
    vN[2 * ii] = -t[1] / len;
    vN[2 * ii + 1] = t[0] / len;
    P0[2 * ii] = a[0];
    P0[2 * ii + 1] = a[1];
  }
  (void)xRelease;
  return 0;
}

/* TR_Find_Shear_Coeffs2 trKernelsShear.cl:414-496 with reorganize_values
 * :380-399 and the Gaussian weight :403-411: the four nearest wall nodes to
 * the midpoint of each shear boundary link, their normalised weights, both
 * sorted by wall-node index. */
void orc_shear_coeffs2(int XF, const int32_t *bounds, const double *P0, const double *vN,
                       const double *blLen, const int32_t *ssArr, const int32_t *ssInds,
                       int ssMinX, int ssMaxX, int irs, double cutoff, double *W, int32_t *S) {
  int nbot = bounds[1], ntot = bounds[1] + bounds[3];
  const double PI = 3.1415926535897931;
  for (int ii = 0; ii < ntot; ii++) {
    int bl = ii < nbot ? ii + bounds[0] : ii + bounds[2] - nbot;
    double half = blLen[bl] / 2.;
    double mid[2] = {P0[2 * bl] + half * vN[2 * bl + 1], P0[2 * bl + 1] + half * (-vN[2 * bl])};
    int ival = (int)mid[0];
    int a = ival - irs > ssMinX ? ival - irs : ssMinX;
    int b = ival + irs + 1 < ssMaxX ? ival + irs + 1 : ssMaxX;
    int start = ssInds[a], stop = ssInds[b];
    int idx[4] = {-1, -1, -1, -1};
    double dr[4] = {100., 100., 100., 100.};
    for (int j = start; j < stop; j++) {
      int h = ssArr[j];
      double dx = mid[0] - (double)(h % XF), dy = mid[1] - (double)(h / XF);
      double dist = sqrt(dx * dx + dy * dy);
      if (dist < dr[0] && dist < cutoff) {
        dr[0] = dist;
        idx[0] = j;
        for (int k = 0; k < 3; k++) { /* keep dr[0] >= dr[1] >= dr[2] >= dr[3] */
          if (dr[k] >= dr[k + 1]) break;
          double td = dr[k]; dr[k] = dr[k + 1]; dr[k + 1] = td;
          int ti = idx[k]; idx[k] = idx[k + 1]; idx[k + 1] = ti;
        }
      }
    }
    double w[4];
    for (int k = 0; k < 4; k++) {
      double s = dr[k] / cutoff;
      w[k] = 1. / sqrt(2. * PI) * exp(-0.5 * s * s);
    }
    static const int net[5][2] = {{0, 1}, {2, 3}, {0, 2}, {1, 3}, {1, 2}};
    for (int k = 0; k < 5; k++) {
      int p = net[k][0], r = net[k][1];
      if (idx[p] > idx[r]) {
        int ti = idx[p]; idx[p] = idx[r]; idx[r] = ti;
        double tw = w[p]; w[p] = w[r]; w[r] = tw;
      }
    }
    double sum = ((w[0] + w[1]) + w[2]) + w[3];
    for (int k = 0; k < 4; k++) {
      W[4 * ii + k] = w[k] / sum;
      S[4 * ii + k] = idx[k];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Lattice Boltzmann                                                         */
/* ------------------------------------------------------------------------ */

/* Product-form equilibrium, clVariablesLB.cpp:77-117:
 * feq_i = rho * phi(cx_i, ux) * phi(cy_i, uy),
 * phi(0,u) = 2/3 - u^2, phi(+-1,u) = (1/3 +- u + u^2)/2. */
static double phi(int c, double u) {
  if (c == 0) return 2. / 3. - u * u;
  return (1. / 3. + c * u + u * u) * 0.5;
}

void orc_feq(double rho, double ux, double uy, double *feq) {
  for (int i = 0; i < 9; i++) feq[i] = rho * phi(CX[i], ux) * phi(CY[i], uy);
}

/* iniDists / iniDistsFromVel (clVariablesLB.cpp:367-425): f = feq(rho, u) on
 * non-solid nodes, 0 elsewhere (iniDists leaves solid nodes at zero). */
void orc_init_dists(int nX, int XF, int nY, const int16_t *nType, const double *rho,
                    const double *ux, const double *uy, double *F) {
  const long D = (long)XF * nY;
  memset(F, 0, sizeof(double) * 9 * D);
  for (int j = 0; j < nY; j++)
    for (int i = 0; i < nX; i++) {
      long g = i + (long)XF * j;
      if (nType[g] & M_SOLID_NODE) continue;
      double feq[9];
      orc_feq(rho[g], ux[g], uy[g], feq);
      for (int k = 0; k < 9; k++) F[g + D * k] = feq[k];
    }
}

typedef struct {
  int nX, nY, XF;
  double tau0;   /* reference 'tau0': relaxation FREQUENCY 1/(3 nu + 0.5) */
  double Fval;   /* FTERM_VAL */
  double nu;     /* MU_NUMBER */
  double maxNut; /* MAX_TURB_VISC */
  double dt;     /* k-omega time step DTFD */
  int options;   /* OPTION_SAVE_MACRO_FIELDS */
  int threads;   /* OpenMP threads for the time loops (0 = runtime default) */
} orc_par;

/* moments in the reference summation order (lbKernels.cl:117-135) */
static void moments(const double *f, double *rho, double *ux, double *uy) {
  double r = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
  double inv = 1. / r;
  *rho = r;
  *ux = inv * (f[1] - f[2] + f[5] - f[6] + f[7] - f[8]);
  *uy = inv * (f[3] - f[4] + f[5] - f[6] - f[7] + f[8]);
}

/* push a collided node: halfway bounce-back into its own node for every
 * *_BOUND link (postCollisionBC lbKernels.cl:166-171), streaming with the
 * x wrap at nX otherwise (lbKernels.cl:306-333). */
static void push_node(const orc_par *p, long D, int gx, long g, int type, const double *f,
                      double *FB) {
  FB[g] = f[0];
  int e = gx < p->nX - 1 ? 1 : -gx;
  int w = gx > 0 ? -1 : p->nX - 1;
  for (int i = 1; i < 9; i++) {
    if (type & BOUND[i]) {
      FB[g + D * REV[i]] = f[i];
    } else {
      long off = (CX[i] > 0 ? e : (CX[i] < 0 ? w : 0)) + (long)CY[i] * p->XF;
      FB[g + off + D * i] = f[i];
    }
  }
}

/* LB_collision_SRT_Fluid lbKernels.cl:175-335 (k-omega off) and
 * LB_collision_SRT_Fluid_w_kOmega :345-633 (k-omega on): read, moments,
 * optional macro output (u_x + F/2), exact-difference forcing
 * f <- f - w (f - feq(u)) + feq(u + F) - feq(u), bounce-back, push.
 * With k-omega: clamp k, omega (:400-410), F2 from the previous strain
 * (:411-417), nu_t (:420), w_t = 1/(3(nu_t+nu)+0.5) (:451), strain from the
 * pre-collision non-equilibrium with the laminar w0 (:518-583). */
void orc_collide_stream(const orc_par *p, const int16_t *nType, const double *FA,
                        double *FB, double *rho_a, double *ux_a, double *uy_a, double *K,
                        double *O, const double *WallD, double *Nut, double *Sxy, int komega) {
  const long D = (long)p->XF * p->nY;
#pragma omp parallel for schedule(static) num_threads(p->threads > 0 ? p->threads : omp_get_max_threads())
  for (int gy = 0; gy < p->nY; gy++) {
    for (int gx = 0; gx < p->nX; gx++) {
      long g = gx + (long)p->XF * gy;
      int type = nType[g];
      if (type & M_SOLID_NODE) continue;
      double f[9], rho, ux, uy;
      for (int i = 0; i < 9; i++) f[i] = FA[g + D * i];
      moments(f, &rho, &ux, &uy);
      double wt = p->tau0;
      if (komega) {
        double om = O[g], k = K[g], y = WallD[g];
        if (om < 0.) {
          om = fabs(om);
          O[g] = om;
        }
        if (k < 0.) {
          k = 1e-16;
          K[g] = 1e-16;
        }
        double a = 2. * sqrt(k) / (KO_BETA_STAR * om * y);
        double b = 500. * p->nu / (y * y * om);
        double arg2 = a > b ? a : b;
        double F2 = tanh(arg2 * arg2);
        double sxx = Sxy[g], sxy = Sxy[g + D], syy = Sxy[g + 2 * D];
        F2 *= sqrt(2. * (sxx * sxx + 2. * sxy * sxy + syy * syy));
        double den = KO_A1 * om > F2 ? KO_A1 * om : F2;
        double nt = KO_A1 * k / den;
        nt = nt < p->maxNut ? nt : p->maxNut;
        nt = nt > 1e-16 ? nt : 1e-16;
        Nut[g] = nt;
        wt = 1. / (3. * (nt + p->nu) + 0.5);
      }
      if (p->options & OPTION_SAVE_MACRO_FIELDS) {
        rho_a[g] = rho;
        ux_a[g] = ux + p->Fval / 2.;
        uy_a[g] = uy;
      }
      double feq[9], feqF[9];
      orc_feq(rho, ux, uy, feq);
      orc_feq(rho, ux + p->Fval, uy, feqF);
      if (komega) {
        double sxx = 0., syy = 0., sxy = 0.;
        for (int i = 1; i < 9; i++) {
          double fn = f[i] - feq[i];
          sxx += CX[i] * CX[i] * fn;
          syy += CY[i] * CY[i] * fn;
          sxy += CX[i] * CY[i] * fn;
        }
        double c = -1.5 * p->tau0 / rho;
        Sxy[g] = c * sxx;
        Sxy[g + D] = c * sxy;
        Sxy[g + 2 * D] = c * syy;
      }
      for (int i = 0; i < 9; i++) f[i] = f[i] - wt * (f[i] - feq[i]) + (feqF[i] - feq[i]);
      push_node(p, D, gx, g, type, f, FB);
    }
  }
}

/* lbIBB lbKernels.cl:640-668 with ibbRev/ibbNeigh Heading.cl:343-356
 * (Bouzidi linear interpolation).  All inputs are read before any output is
 * written; a link listed twice takes its last q. */
void orc_ibb(const orc_par *p, const int32_t *ibb, const double *q, int n, double *FB) {
  const long D = (long)p->XF * p->nY;
  double *nv = (double *)malloc(sizeof(double) * (n > 0 ? n : 1));
  for (int i = 0; i < n; i++) {
    long loc = ibb[i];
    int dir = (int)(loc / D);
    long g = loc - D * dir;
    long rd = g + D * REV[dir];
    double tq = 2. * q[i];
    if (tq <= 1.0) {
      nv[i] = tq * FB[rd] + (1. - tq) * FB[loc];
    } else {
      int gx = (int)(g % p->XF);
      int sx = gx - CX[dir];
      sx = sx >= p->nX ? sx - p->nX : (sx < 0 ? sx + p->nX : sx);
      long nb = sx + (g / p->XF - CY[dir]) * (long)p->XF + D * REV[dir];
      nv[i] = 1. / tq * (FB[rd] + (tq - 1.) * FB[nb]);
    }
  }
  for (int i = 0; i < n; i++) {
    long loc = ibb[i];
    int dir = (int)(loc / D);
    FB[loc - D * dir + D * REV[dir]] = nv[i];
  }
  free(nv);
}

/* ------------------------------------------------------------------------ */
/* k-omega                                                                   */
/* ------------------------------------------------------------------------ */

/* Nonuniform three-point first-derivative coefficients and second-derivative
 * coefficients from the wall-clamped spacings (kOmegaKernels.cl:37-55,
 * 235-257). c = {e, w, c, e2, w2, c2} along one axis. */
void orc_fd_coeffs(double dp, double dm, double *c) {
  dp = dp > 0.1 ? dp : 0.1;
  dm = dm > 0.1 ? dm : 0.1;
  double dc = dp + dm;
  dc = dc > 0.1 ? dc : 0.1;
  c[0] = dm / (dp * dc);
  c[1] = -dp / (dm * dc);
  c[2] = (dp - dm) / (dp * dm);
  if (fabs(c[2]) > 10.) {
    c[0] = 1.0 / dc;
    c[1] = -1.0 / dc;
    c[2] = 0.0;
  }
  c[3] = 2. / (dp * dc);
  c[4] = 2. / (dm * dc);
  c[5] = -2. / (dp * dm);
}

/* Update_kOmega_Diffusivities kOmegaKernels.cl:7-95: F1 blending, D_k, D_w
 * and the cross-diffusion vector 2 sigma_w2 / w * grad k. */
void orc_ko_diffusivities(const orc_par *p, const int16_t *nType, const double *K,
                          const double *O, const double *Nut, const double *WallD,
                          double *Dk, double *Do, double *F1a, double *dkdo, const double *rho,
                          const double *dX) {
  const long D = (long)p->XF * p->nY;
#pragma omp parallel for schedule(static) num_threads(p->threads > 0 ? p->threads : omp_get_max_threads())
  for (int j = 0; j < p->nY; j++)
    for (int i = 0; i < p->nX; i++) {
      long g = i + (long)p->XF * j;
      if (nType[g] & M_SOLID_NODE) continue;
      double cx[6], cy[6];
      orc_fd_coeffs(dX[g], dX[g + D], cx);
      orc_fd_coeffs(dX[g + 2 * D], dX[g + 3 * D], cy);
      long gw = i > 0 ? g - 1 : p->nX - 1 + (long)j * p->XF;
      long ge = i < p->nX - 1 ? g + 1 : (long)j * p->XF;
      double om = O[g], k = K[g], y = WallD[g];
      double dOdx = cx[0] * O[ge] + cx[1] * O[gw] + cx[2] * om;
      double dKdx = cx[0] * K[ge] + cx[1] * K[gw] + cx[2] * k;
      double dOdy = cy[0] * O[g + p->XF] + cy[1] * O[g - p->XF] + cy[2] * om;
      double dKdy = cy[0] * K[g + p->XF] + cy[1] * K[g - p->XF] + cy[2] * k;
      double m = 2. * KO_SIGMA_O2 / om;
      double mx = m * dKdx, my = m * dKdy;
      dkdo[g] = mx;
      dkdo[g + D] = my;
      double cd = rho[g] * (mx * dOdx + my * dOdy);
      cd = cd > 1.0e-10 ? cd : 1.0e-10;
      double a = sqrt(k) / (KO_BETA_STAR * om * y), b = 500. * p->nu / (y * y * om);
      double arg = a > b ? a : b;
      double c = 4. * k * KO_SIGMA_O2 / (cd * y * y);
      arg = arg < c ? arg : c;
      double F1 = tanh(arg * arg * arg * arg);
      F1a[g] = F1;
      double nt = Nut[g];
      Dk[g] = p->nu + (KO_SIGMA_K1 * F1 + KO_SIGMA_K2 * (1. - F1)) * nt;
      Do[g] = p->nu + (KO_SIGMA_O1 * F1 + KO_SIGMA_O2 * (1. - F1)) * nt;
    }
}

/* Update_kOmega_Coeffs_Implicit kOmegaKernels.cl:187-480: diagonal-normalised
 * five-point rows.  A[g + D*m], m = 0..3 for E, W, N, S (the diagonal is 1),
 * b[g].  NESW-boundary fluid nodes get the omega wall row (:389-403). */
void orc_ko_coeffs(const orc_par *p, const int16_t *nType, const double *K, const double *O,
                   const double *Nut, const double *Dk, const double *Do, const double *F1a,
                   const double *dkdo, const double *Sxy, const double *rho, const double *ux,
                   const double *uy, const double *dX, const double *WallD, double *Ak,
                   double *Ao, double *bk, double *bo) {
  const long D = (long)p->XF * p->nY;
#pragma omp parallel for schedule(static) num_threads(p->threads > 0 ? p->threads : omp_get_max_threads())
  for (int j = 0; j < p->nY; j++)
    for (int i = 0; i < p->nX; i++) {
      long g = i + (long)p->XF * j;
      int t = nType[g];
      if (t & M_SOLID_NODE) continue;
      double cx[6], cy[6];
      orc_fd_coeffs(dX[g], dX[g + D], cx);
      orc_fd_coeffs(dX[g + 2 * D], dX[g + 3 * D], cy);
      long gw = i > 0 ? g - 1 : p->nX - 1 + (long)j * p->XF;
      long ge = i < p->nX - 1 ? g + 1 : (long)j * p->XF;
      long gn = g + p->XF, gs = g - p->XF;
      double dk = Dk[g], dw = Do[g];
      double dkdx = cx[0] * Dk[ge] + cx[1] * Dk[gw] + cx[2] * dk;
      double dwdx = cx[0] * Do[ge] + cx[1] * Do[gw] + cx[2] * dw;
      double dkdy = cy[0] * Dk[gn] + cy[1] * Dk[gs] + cy[2] * dk;
      double dwdy = cy[0] * Do[gn] + cy[1] * Do[gs] + cy[2] * dw;
      double Ux = ux[g], Uy = uy[g], R = rho[g];
      double Jxw = Ux - dwdx - dkdo[g], Jyw = Uy - dwdy - dkdo[g + D];
      double Jxk = Ux - dkdx, Jyk = Uy - dkdy;
      double om = O[g] > 0. ? O[g] : 0., k = K[g] > 0. ? K[g] : 0.;
      double F1 = F1a[g], nt = Nut[g];
      double sxx = Sxy[g], sxy = Sxy[g + D], syy = Sxy[g + 2 * D];
      double Pk = 2. * nt * (sxx * sxx + syy * syy + 2. * sxy * sxy) -
                  2. / 3. * k * R * (sxx + syy);
      double lim = 10. * KO_BETA_STAR * k * om;
      Pk = Pk < lim ? Pk : lim;
      Pk = Pk > 0. ? Pk : 0.;
      double gam = (F1 * KO_GAMMA1 + (1. - F1) * KO_GAMMA2) / nt;
      double Sc = -KO_BETA_STAR * om;
      double dt = p->dt;
      double Kc = 1. + dt * (Jxk * cx[2] + Jyk * cy[2] - dk * (cx[5] + cy[5])) - Sc;
      double inv = 1. / Kc;
      Ak[g] = dt * (Jxk * cx[0] - dk * cx[3]) * inv;
      Ak[g + D] = dt * (Jxk * cx[1] - dk * cx[4]) * inv;
      Ak[g + 2 * D] = dt * (Jyk * cy[0] - dk * cy[3]) * inv;
      Ak[g + 3 * D] = dt * (Jyk * cy[1] - dk * cy[4]) * inv;
      bk[g] = (k + dt * Pk) * inv;
      if (t & NESW_BOUNDARY_NODE) {
        dt = 0.;
        Sc = 0.;
        double yp = WallD[g];
        double ov = (6. * p->nu / 0.075) / (yp * yp);
        double ol = sqrt(k) / (KO_C_MU_0_25 * KO_KAPPA * yp);
        double ow = sqrt(ov * ov + ol * ol);
        om = ow < 1000. ? ow : 1000.;
      }
      double Wc = 1. + dt * (Jxw * cx[2] + Jyw * cy[2] - dw * (cx[5] + cy[5])) - Sc;
      inv = 1. / Wc;
      Ao[g] = dt * (Jxw * cx[0] - dw * cx[3]) * inv;
      Ao[g + D] = dt * (Jxw * cx[1] - dw * cx[4]) * inv;
      Ao[g + 2 * D] = dt * (Jyw * cy[0] - dw * cy[3]) * inv;
      Ao[g + 3 * D] = dt * (Jyw * cy[1] - dw * cy[4]) * inv;
      bo[g] = (om + dt * (gam * Pk)) * inv;
    }
}

/* y = A x on fluid rows (diagonal 1, four neighbours with the x wrap at nX);
 * other rows give y = x, which keeps solid-boundary entries (x = 0) fixed. */
static void ko_apply(const orc_par *p, const int16_t *nType, int solid, const double *A,
                     const double *x, double *y) {
  const long D = (long)p->XF * p->nY;
  for (long g = 0; g < D; g++) y[g] = x[g];
  for (int j = 0; j < p->nY; j++)
    for (int i = 0; i < p->nX; i++) {
      long g = i + (long)p->XF * j;
      if (nType[g] & solid) continue;
      long gw = i > 0 ? g - 1 : p->nX - 1 + (long)j * p->XF;
      long ge = i < p->nX - 1 ? g + 1 : (long)j * p->XF;
      y[g] = x[g] + (A[g] * x[ge] + A[g + D] * x[gw] + A[g + 2 * D] * x[g + p->XF] +
                     A[g + 3 * D] * x[g - p->XF]);
    }
}

static double fluid_dot(const orc_par *p, const int16_t *nType, int solid, const double *a,
                        const double *b) {
  double s = 0.;
  for (int j = 0; j < p->nY; j++)
    for (int i = 0; i < p->nX; i++) {
      long g = i + (long)p->XF * j;
      if (!(nType[g] & solid)) s += a[g] * b[g];
    }
  return s;
}

/* The preconditioner's pairing of node (i, j): +1 paired with the node above,
 * -1 with the node below, 0 unpaired.  In each column the first two fluid
 * nodes above a solid stretch form a pair, and the last two below one unless
 * the lower of them is already paired (nodes outside the lattice count as
 * solid). */
static int pair_dir(const orc_par *p, const int16_t *nType, int solid, int i, int j) {
  if (nType[i + (long)p->XF * j] & solid) return 0;
#define ORC_S(k) (j - (k) < 0 || (nType[i + (long)p->XF * (j - (k))] & solid))
#define ORC_N(k) (j + (k) >= p->nY || (nType[i + (long)p->XF * (j + (k))] & solid))
  int d = 0;
  if (ORC_S(1) && !ORC_N(1))
    d = 1;
  else if (!ORC_S(1) && ORC_S(2))
    d = -1;
  else if (ORC_N(1) && !ORC_S(1) && !ORC_S(2) && !ORC_S(3))
    d = -1;
  else if (!ORC_N(1) && ORC_N(2) && !ORC_S(1) && !ORC_S(2))
    d = 1;
#undef ORC_S
#undef ORC_N
  return d;
}

/* z = M^-1 u: M is A's 2 x 2 blocks over the pairs (identity elsewhere), so
 * z_g = (u_g - a_gq u_q) / (1 - a_gq a_qg) with q the partner. */
static void pair_apply(const orc_par *p, const int16_t *nType, int solid, const int8_t *pr,
                       const double *A, const double *u, double *z) {
  const long D = (long)p->XF * p->nY;
  for (long g = 0; g < D; g++) {
    if (nType[g] & solid || (g % p->XF) >= p->nX) {
      z[g] = u[g];
      continue;
    }
    if (!pr[g]) {
      z[g] = u[g];
      continue;
    }
    long q = pr[g] > 0 ? g + p->XF : g - p->XF;
    double agq = A[g + (pr[g] > 0 ? 2 : 3) * D], aqg = A[q + (pr[g] > 0 ? 3 : 2) * D];
    z[g] = (u[g] - agq * u[q]) / (1. - agq * aqg);
  }
}

/* Solves x_g + sum_m A[g+D*m] x_nb(m) = b_g on fluid rows (solid-boundary rows
 * are identity with b = 0, GMRESSolver.cpp:1421-1433, so x = 0 there; interior
 * solid rows are empty) from the current x, with BiCGStab (van der Vorst
 * 1992) right-preconditioned by the wall pairs' 2 x 2 blocks, in FP64, until
 * ||r|| <= rtol ||r0|| or ||r|| <= atol (the criterion of
 * GMRESSolver::checkConvergence, GMRESSolver.cpp:1535-1560), or once the
 * residual has stopped falling for 50 iterations (rounding floor).  The
 * reference runs GMRES(10) (ConstDef.h:20-21) without a preconditioner;
 * converged, both give the solution of the same system.  Returns the
 * iteration count, or -1 on a non-finite residual (the reference halves its
 * time step then).  work: 8*D doubles plus D bytes. */
/* mode 0: GMRESSolver::checkConvergence (GMRESSolver.cpp:1535-1560):
 *   ||r|| <= rtol ||r0|| or ||r|| <= atol, non-finite residual fails;
 * mode 1: BiCGStabSolver (BiCGStabSolver.cpp:509-545, 628-660): with
 *   e = ||r|| / ||b||, e <= rtol or ||r|| <= atol ||r0||; ||b|| <= 1e-16 gives
 *   x = b; a non-finite e, or e above 10 times the previous iteration's after
 *   the first, fails (the solver then halves the time step). */
static int solve5(const orc_par *p, const int16_t *nType, int solid, int mode, const double *A,
                  const double *b, double *x, double rtol, double atol, int maxit, double *res0,
                  double *res, double *work) {
  const long D = (long)p->XF * p->nY;
  double *r = work, *rh = work + D, *pp = work + 2 * D, *v = work + 3 * D, *s = work + 4 * D,
         *t = work + 5 * D, *ph = work + 6 * D, *sh = work + 7 * D;
  int8_t *pr = (int8_t *)(work + 8 * D);
  for (int j = 0; j < p->nY; j++)
    for (int i = 0; i < p->XF; i++) pr[i + (long)p->XF * j] = i < p->nX ? pair_dir(p, nType, solid, i, j) : 0;
  double bn = 0.;
  if (mode == 1) {
    bn = sqrt(fluid_dot(p, nType, solid, b, b));
    if (bn <= 1.0e-16) {
      for (long g = 0; g < D; g++)
        if (!(nType[g] & solid) && (g % p->XF) < p->nX) x[g] = b[g];
      return 0;
    }
  }
  ko_apply(p, nType, solid, A, x, t);
  for (long g = 0; g < D; g++) {
    int fl = !(nType[g] & solid) && (g % p->XF) < p->nX;
    r[g] = fl ? b[g] - t[g] : 0.;
    rh[g] = r[g];
    pp[g] = 0.;
    v[g] = 0.;
  }
  double r0 = sqrt(fluid_dot(p, nType, solid, r, r));
  if (res0) *res0 = r0;
  if (res) *res = r0;
  if (!isfinite(r0)) return -1;
  if (mode == 0 ? (r0 <= atol || r0 == 0.) : (r0 <= rtol * bn || r0 == 0.)) return 0;
  double prev = r0;
  double rho = 1., alpha = 1., om = 1., best = r0;
  int since_best = 0;
  for (int it = 1; it <= maxit; it++) {
    double rho1 = fluid_dot(p, nType, solid, rh, r);
    if (rho1 == 0.) { /* breakdown: restart from the current residual */
      for (long g = 0; g < D; g++) { rh[g] = r[g]; pp[g] = 0.; v[g] = 0.; }
      rho = alpha = om = 1.;
      rho1 = fluid_dot(p, nType, solid, rh, r);
    }
    double beta = (rho1 / rho) * (alpha / om);
    for (long g = 0; g < D; g++) pp[g] = r[g] + beta * (pp[g] - om * v[g]);
    pair_apply(p, nType, solid, pr, A, pp, ph);
    ko_apply(p, nType, solid, A, ph, v);
    alpha = rho1 / fluid_dot(p, nType, solid, rh, v);
    for (long g = 0; g < D; g++) s[g] = r[g] - alpha * v[g];
    pair_apply(p, nType, solid, pr, A, s, sh);
    ko_apply(p, nType, solid, A, sh, t);
    double tt = fluid_dot(p, nType, solid, t, t);
    om = tt > 0. ? fluid_dot(p, nType, solid, t, s) / tt : 0.;
    for (long g = 0; g < D; g++) {
      int fl = !(nType[g] & solid) && (g % p->XF) < p->nX;
      if (!fl) continue;
      x[g] += alpha * ph[g] + om * sh[g];
      r[g] = s[g] - om * t[g];
    }
    rho = rho1;
    double rn = sqrt(fluid_dot(p, nType, solid, r, r));
    if (res) *res = rn;
    if (!isfinite(rn)) return -1;
    if (mode == 0) {
      if (rn <= rtol * r0 || rn <= atol) return it;
    } else {
      if (it > 1 && rn > 10. * prev) return -1;
      if (rn <= rtol * bn || rn <= atol * r0) return it;
      prev = rn;
    }
    /* converged to the rounding floor: no progress in 50 iterations */
    if (rn < 0.999 * best) {
      best = rn;
      since_best = 0;
    } else if (++since_best >= 50) {
      return it;
    }
    if (om == 0.) { /* stagnation: restart */
      for (long g = 0; g < D; g++) { rh[g] = r[g]; pp[g] = 0.; v[g] = 0.; }
      rho = alpha = om = 1.;
    }
  }
  return maxit;
}

/* The k and omega systems: rows on the nodes that are not solid. */
int orc_ko_solve(const orc_par *p, const int16_t *nType, const double *A, const double *b,
                 double *x, double rtol, double atol, int maxit, double *res0, double *res,
                 double *work) {
  return solve5(p, nType, M_SOLID_NODE, 0, A, b, x, rtol, atol, maxit, res0, res, work);
}

/* The temperature system: rows on the nodes that are not solid in the clean
 * geometry (fluid and the fouling layer, clVariablesFD's CSR_Inds built on
 * the M0 flags), with BiCGStabSolver's stopping rule. */
int orc_t_solve(const orc_par *p, const int16_t *nType, const double *A, const double *b,
                double *x, double rtol, double atol, int maxit, double *res0, double *res,
                double *work) {
  return solve5(p, nType, M0_SOLID_NODE, 1, A, b, x, rtol, atol, maxit, res0, res, work);
}

/* ------------------------------------------------------------------------ */
/* Wall shear                                                                */
/* ------------------------------------------------------------------------ */

/* TR_shear_1 trKernelsShear.cl:174-222 with calcLBDists :131-171: signed
 * traction magnitude at each wall node from the newest populations and the
 * stored macros (u_x includes +F/2). */
void orc_shear1(const orc_par *p, const int32_t *ssArr, int nSS, const double *CFSS,
                const double *F, const double *rho, const double *ux, const double *uy,
                const double *Nut, double *Tau) {
  const long D = (long)p->XF * p->nY;
  for (int ii = 0; ii < nSS; ii++) {
    long g = ssArr[ii];
    double feq[9];
    orc_feq(rho[g], ux[g], uy[g], feq);
    double ve = (Nut ? Nut[g] : 0.) + p->nu;
    ve *= 1. / (3. * ve + 0.5);
    double tx = 0., ty = 0.;
    for (int a = 0; a < 8; a++) {
      double fn = F[g + D * (a + 1)] - feq[a + 1];
      tx += fn * (CFSS[16 * ii + a] * ve);
      ty += fn * (CFSS[16 * ii + 8 + a] * ve);
    }
    double t = sqrt(tx * tx + ty * ty) / rho[g];
    Tau[ii] = tx < 0 ? -t : t;
  }
}

/* TR_shear_2 trKernelsShear.cl:300-316: link shear from four wall nodes.
 * blTau is indexed by the shear boundary-link index (GET_SHEAR_BL_INDEX). */
void orc_shear2(const int32_t *bounds, const double *W, const int32_t *S, const double *Tau,
                double *blTau) {
  int ntot = bounds[1] + bounds[3];
  for (int ii = 0; ii < ntot; ii++) {
    double s = 0.;
    for (int k = 0; k < 4; k++) s += W[4 * ii + k] * (S[4 * ii + k] >= 0 ? Tau[S[4 * ii + k]] : 0.);
    blTau[ii] = s;
  }
}

/* ------------------------------------------------------------------------ */
/* Reductions                                                                */
/* ------------------------------------------------------------------------ */

/* Reducer over the zero-padded power-of-two buffer (Reducer.h:88-130,
 * ReduceGenerator.cpp:219-305; FullSize clProblem.cpp:476-480): op 1 sum,
 * 2 max, 4 min.  Padding contributes 0 (so a padded min is <= 0). */
double orc_reduce(const double *a, long n, int op) {
  long full = 2;
  while (full < n) full *= 2;
  double acc = op == 1 ? 0. : a[0];
  for (long i = 0; i < n; i++) {
    if (op == 1)
      acc += a[i];
    else if (op == 2)
      acc = a[i] > acc ? a[i] : acc;
    else
      acc = a[i] < acc ? a[i] : acc;
  }
  if (full > n) {
    if (op == 2) acc = acc > 0. ? acc : 0.;
    if (op == 4) acc = acc < 0. ? acc : 0.;
  }
  return acc;
}

/* ------------------------------------------------------------------------ */
/* Whole steps                                                               */
/* ------------------------------------------------------------------------ */

/* Solver settings and the k-omega time step control state (kOmega.h:22,
 * 204-205): sub-steps per LB step, their time step and the steps since the
 * last change.  inject_nan makes the next solve report a non-finite residual
 * (a test of the restart path without a non-finite value in any array);
 * halvings counts the restarts. */
typedef struct {
  double rtol, atol;
  int maxit;
  int last_iters_k, last_iters_o;
  int steps_per_lb;
  double dt_cur;
  int since_change;
  int inject_nan;
  int halvings;
} orc_solver;

#define ORC_DT_QUIET 20          /* kOmega.h:22 timeBtwIncreaseTimeStep */
#define ORC_MAX_STEPS_PER_LB (1 << 16)

/* One reference time step (clVariablesLB::Solve clVariablesLB.h:375-441):
 * collide+stream FA->FB, IBB on FB, then (k-omega on) kOmega::Solve
 * (kOmega.cpp:31-73): the time step doubles back after ORC_DT_QUIET steps
 * without a change (doubleTimeStep :401-414), k and omega are kept
 * (copyToPrevSolution), and each of the steps_per_lb sub-steps runs the
 * diffusivities, the coefficients with the current time step, the omega
 * solve and the k solve.  A non-finite residual restores k and omega, halves
 * the time step and doubles the sub-steps (halveTimeStep :387-399 via
 * GMRESSolver::checkConvergence, GMRESSolver.cpp:1535-1546); the sub-steps
 * then start again from the first.  (The reference resumes its loop at the
 * second sub-step after running the k solve on rows assembled with the old
 * time step; here the restart is clean.)  The caller swaps FA/FB.  work:
 * 26*D doubles. */
int orc_step(const orc_par *p, const int16_t *nType, const int32_t *ibb, const double *q,
             int nIbb, const double *FA, double *FB, double *rho, double *ux, double *uy,
             double *K, double *O, const double *WallD, double *Nut, double *Sxy,
             const double *dX, int komega, orc_solver *sv, double *work) {
  const long D = (long)p->XF * p->nY;
  orc_collide_stream(p, nType, FA, FB, rho, ux, uy, K, O, WallD, Nut, Sxy, komega);
  orc_ibb(p, ibb, q, nIbb, FB);
  if (!komega) return 0;
  double *Dk = work, *Do = work + D, *F1 = work + 2 * D, *dkdo = work + 3 * D;
  double *Ak = work + 5 * D, *Ao = Ak + 4 * D, *bk = Ao + 4 * D, *bo = bk + D;
  double *sw = bo + D; /* solver work, 8*D doubles and D bytes */
  double *Kp = sw + 9 * D, *Op = Kp + D; /* restart point: work must hold 26*D doubles */
  if (sv->steps_per_lb < 1) {
    sv->steps_per_lb = 1;
    sv->dt_cur = p->dt;
  }
  if (++sv->since_change >= ORC_DT_QUIET) {
    sv->since_change = 0;
    if (sv->steps_per_lb > 1) {
      sv->steps_per_lb /= 2;
      sv->dt_cur *= 2.;
    }
  }
  memcpy(Kp, K, sizeof(double) * D);
  memcpy(Op, O, sizeof(double) * D);
  orc_par pp = *p;
  for (int sub = 0; sub < sv->steps_per_lb;) {
    pp.dt = sv->dt_cur;
    orc_ko_diffusivities(&pp, nType, K, O, Nut, WallD, Dk, Do, F1, dkdo, rho, dX);
    orc_ko_coeffs(&pp, nType, K, O, Nut, Dk, Do, F1, dkdo, Sxy, rho, ux, uy, dX, WallD, Ak, Ao,
                  bk, bo);
    int io = orc_ko_solve(&pp, nType, Ao, bo, O, sv->rtol, sv->atol, sv->maxit, 0, 0, sw);
    int ik = io < 0 ? -1 : orc_ko_solve(&pp, nType, Ak, bk, K, sv->rtol, sv->atol, sv->maxit, 0, 0, sw);
    if (sv->inject_nan) {
      sv->inject_nan = 0;
      io = -1;
    }
    sv->last_iters_o = io;
    sv->last_iters_k = ik;
    if (io < 0 || ik < 0) {
      if (sv->steps_per_lb >= ORC_MAX_STEPS_PER_LB) return -1;
      memcpy(K, Kp, sizeof(double) * D);
      memcpy(O, Op, sizeof(double) * D);
      sv->steps_per_lb *= 2;
      sv->dt_cur /= 2.;
      sv->since_change = 0;
      sv->halvings++;
      sub = 0;
      continue;
    }
    sub++;
  }
  return 0;
}

/* n steps with FA/FB ping-pong; on return the newest populations are in FA. */
int orc_run(const orc_par *p, const int16_t *nType, const int32_t *ibb, const double *q,
            int nIbb, double *FA, double *FB, double *rho, double *ux, double *uy, double *K,
            double *O, const double *WallD, double *Nut, double *Sxy, const double *dX,
            int komega, orc_solver *sv, double *work, int nsteps) {
  const long D = (long)p->XF * p->nY;
  double *a = FA, *b = FB;
  int rc = 0;
  for (int s = 0; s < nsteps && !rc; s++) {
    rc = orc_step(p, nType, ibb, q, nIbb, a, b, rho, ux, uy, K, O, WallD, Nut, Sxy, dX, komega,
                  sv, work);
    double *t = a;
    a = b;
    b = t;
  }
  if (a != FA) memcpy(FA, a, sizeof(double) * 9 * D);
  return rc;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Temperature (clVariablesFD)                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
  double alpha_fluid; /* Alpha_fluid = MuVal / Pr (clVariablesFD.cpp:458) */
  double alpha_foul;  /* Alpha_foul, fouling-layer nodes (:463) */
  double pr_turb_inv; /* 1 / PrTurbNum (PR_TURB_NUMBER_INV) */
  double t_in;        /* TFD_X_IN_VAL, the inlet temperature "Tinlet Sim" */
  double rtol, atol;  /* "Temp Max Rel Tol", "Temp Max Abs Tol" */
  int maxit;          /* "Temp Max Iterations" */
  int steps_per_lb;   /* stepsPerLB = (int)(1 / dTtfd) */
  double dt;          /* timeStep */
  int since_change;   /* timeSinceTimeStepChange */
  int last_iters, halvings, inject_nan;
} orc_tsolver;

/* One row of the nonuniform three-point differences on the clean-wall
 * spacings, as Update_T_Coeffs_Implicit forms them (tfdKernels.cl:154-181):
 * spacings held at 0.1 or more, the x first derivative falling back to the
 * two-point central difference where |its centre weight| > 10, and the y
 * first derivative where |its north weight| > 10 (the kernel tests dYn). */
static void t_fd(double dp, double dm, int clamp, int test_c, double *d1, double *d2) {
  if (clamp) {
    dp = dp > 0.1 ? dp : 0.1;
    dm = dm > 0.1 ? dm : 0.1;
  }
  double dc = dp + dm;
  if (clamp) dc = dc > 0.1 ? dc : 0.1;
  d1[0] = dm / (dp * dc);          /* +: east / north */
  d1[1] = -dp / (dm * dc);         /* -: west / south */
  d1[2] = (dp - dm) / (dp * dm);   /* centre */
  d2[0] = 2. / (dp * dc);
  d2[1] = 2. / (dm * dc);
  d2[2] = -2. / (dp * dm);
  if (clamp && fabs(test_c ? d1[2] : d1[0]) > 10.) {
    d1[0] = 1. / dc;
    d1[1] = -1. / dc;
    d1[2] = 0.;
  }
}

/* Update_T_Coeffs_Implicit (tfdKernels.cl:122-280) in the five-point row form
 * of the k-omega systems (diagonal 1; planes E, W, N, S): rows on the nodes
 * that are not solid in the clean geometry; a fouling-layer node (solid now,
 * fluid in the clean geometry) takes alpha_foul; nu_t / Pr_t adds to alpha and
 * the x and y differences of nu_t to the advection (the kernel multiplies the
 * centre term by 1 / Pr_t twice, which is kept); the source is T clamped to
 * [0, T_in]; the inlet column's W coefficient moves to the right-hand side
 * with T_in and the outlet column's E coefficient into the diagonal (zero
 * gradient). */
void orc_t_coeffs(const orc_par *p, const int16_t *nType, const double *dX0, const double *T,
                  const double *Nut, const double *ux, const double *uy, const orc_tsolver *tp,
                  double dt, double *A, double *b) {
  const long D = (long)p->XF * p->nY;
  for (int j = 0; j < p->nY; j++)
    for (int i = 0; i < p->nX; i++) {
      long g = i + (long)p->XF * j;
      int t = nType[g];
      if (t & M0_SOLID_NODE) continue;
      double dx1[3], dx2[3], dy1[3], dy2[3];
      t_fd(dX0[g], dX0[g + D], 1, 1, dx1, dx2);
      t_fd(dX0[g + 2 * D], dX0[g + 3 * D], 1, 0, dy1, dy2);
      long gw = i > 0 ? g - 1 : p->nX - 1 + (long)j * p->XF;
      long ge = i < p->nX - 1 ? g + 1 : (long)j * p->XF;
      double alpha = (t & M_SOLID_NODE) ? tp->alpha_foul : tp->alpha_fluid;
      double at = Nut[g] * tp->pr_turb_inv;
      double Jx = ux[g] - tp->pr_turb_inv * (dx1[0] * Nut[ge] + dx1[1] * Nut[gw] + dx1[2] * at);
      double Jy = uy[g] - tp->pr_turb_inv * (dy1[0] * Nut[g + p->XF] + dy1[1] * Nut[g - p->XF] +
                                             dy1[2] * at);
      alpha += at;
      double src = T[g] > 0. ? T[g] : 0.;
      src = src < tp->t_in ? src : tp->t_in;
      double C = 1. + dt * (Jx * dx1[2] + Jy * dy1[2] - alpha * (dx2[2] + dy2[2]));
      double S = dt * (Jy * dy1[1] - alpha * dy2[1]);
      double W = dt * (Jx * dx1[1] - alpha * dx2[1]);
      if (i == 0) src -= tp->t_in * W;
      double E = dt * (Jx * dx1[0] - alpha * dx2[0]);
      if (i == p->nX - 1) C += E;
      double N = dt * (Jy * dy1[0] - alpha * dy2[0]);
      double inv = 1. / C;
      A[g] = i < p->nX - 1 ? E * inv : 0.;
      A[g + D] = i > 0 ? W * inv : 0.;
      A[g + 2 * D] = N * inv;
      A[g + 3 * D] = S * inv;
      b[g] = src * inv;
    }
}

#define ORC_T_QUIET 20 /* clVariablesFD.h:24 timeBtwIncreaseTimeStep */

/* clVariablesFD::Solve (clVariablesFD.h:343-358) with halveTimeStep and
 * doubleTimeStep (clVariablesFD.cpp:202-228): after 20 calls without a change
 * the time step doubles back (never past one sub-step); the start is kept;
 * each sub-step assembles the rows and solves; a failed solve restores the
 * start, halves the time step, doubles the sub-steps and starts them again.
 * work: 4*D + 9*D doubles plus D bytes (rows, right-hand side, start, solver). */
int orc_t_step(const orc_par *p, const int16_t *nType, const double *dX0, double *T,
               const double *Nut, const double *ux, const double *uy, orc_tsolver *tp,
               double *work) {
  const long D = (long)p->XF * p->nY;
  double *A = work, *b = work + 4 * D, *Tp = work + 5 * D, *sw = work + 6 * D;
  if (++tp->since_change >= ORC_T_QUIET) {
    tp->since_change = 0;
    if (tp->steps_per_lb > 1) {
      tp->steps_per_lb /= 2;
      tp->dt *= 2.;
    }
  }
  memcpy(Tp, T, sizeof(double) * D);
  int sub = 0;
  while (sub < tp->steps_per_lb) {
    orc_t_coeffs(p, nType, dX0, T, Nut, ux, uy, tp, tp->dt, A, b);
    int it = orc_t_solve(p, nType, A, b, T, tp->rtol, tp->atol, tp->maxit, 0, 0, sw);
    if (tp->inject_nan) {
      tp->inject_nan = 0;
      it = -1;
    }
    if (it < 0) {
      if (tp->steps_per_lb >= ORC_MAX_STEPS_PER_LB) return -2;
      tp->steps_per_lb *= 2;
      tp->dt /= 2.;
      tp->since_change = 0;
      tp->halvings++;
      memcpy(T, Tp, sizeof(double) * D);
      sub = 0;
      continue;
    }
    tp->last_iters = it;
    sub++;
  }
  return 0;
}

/* ------------------------------------------------------------------------
 * Nusselt number (clVariablesFD::iniNuCoeffs / updateTimeData,
 * clVariablesFD.cpp:362-442, 834-869; tfdKernels.cl:737-1033).
 * ---------------------------------------------------------------------- */

#define NU_EPSILON 1.11e-16 /* Heading.cl:21 */
#define NU_NONE (-1000.)
#define WORKGROUPSIZE_NU 256 /* ConstDef.h:87 */

/* nuFindIntersectionNormal (tfdKernels.cl:737-756). */
static int nu_foot(double Lx, double Ly, double P0x, double P0y, double P1x, double P1y,
                   double *dist, double *Cx, double *Cy) {
  double tx = P1x - P0x, ty = P1y - P0y;
  double nmx = -ty, nmy = tx;
  double nl = sqrt(nmx * nmx + nmy * nmy);
  double nx = nmx / nl, ny = nmy / nl;
  *dist = (Lx - P0x) * nx + (Ly - P0y) * ny;
  *Cx = Lx - nx * *dist;
  *Cy = Ly - ny * *dist;
  double d0x = *Cx - P0x, d0y = *Cy - P0y, d1x = *Cx - P1x, d1y = *Cy - P1y;
  double l10 = sqrt(tx * tx + ty * ty);
  double l0 = sqrt(d0x * d0x + d0y * d0y);
  double l1 = sqrt(d1x * d1x + d1y * d1y);
  return fabs(l10 - l0 - l1) < NU_EPSILON;
}

/* nuFindNearestNode (tfdKernels.cl:759-807, clVariablesFD.cpp:498-536) for
 * link l: the non-solid nodes of its box closer along its normal than the
 * slot's best so far.  Rows 0 and nY - 1 are left out (their getTandDT
 * neighbour would lie outside the lattice). */
static void nu_nearest(int nY, int XF, const int16_t *nType, const double *C, const int32_t *BL,
                       int l, int xStart, int xStop, int shift, double *best, double *vec,
                       int32_t *rows) {
  double P0x = C[2 * BL[2 * l]], P0y = C[2 * BL[2 * l] + 1];
  double P1x = C[2 * BL[2 * l + 1]], P1y = C[2 * BL[2 * l + 1] + 1];
  int x0 = (int)fmin(P0x, P1x) - 1, y0 = (int)fmin(P0y, P1y) - 1;
  int x1 = (int)ceil(fmax(P0x, P1x)) + 1, y1 = (int)ceil(fmax(P0y, P1y)) + 1;
  if (x0 < xStart) x0 = xStart;
  if (y0 < 1) y0 = 1;
  if (x1 > xStop - 1) x1 = xStop - 1;
  if (y1 > nY - 2) y1 = nY - 2;
  if (x0 >= xStop || x1 < xStart) return;
  for (int ii = x0; ii <= x1; ii++)
    for (int jj = y0; jj <= y1; jj++) {
      if (nType[ii + (long)XF * jj] & M_SOLID_NODE) continue;
      double d, cx, cy;
      if (nu_foot(ii, jj, P0x, P0y, P1x, P1y, &d, &cx, &cy) && d < best[ii - shift]) {
        best[ii - shift] = d;
        vec[2 * (ii - shift)] = ii - cx;
        vec[2 * (ii - shift) + 1] = jj - cy;
        rows[ii - shift] = jj;
      }
    }
}

/* getTandDT (tfdKernels.cl:878-917). */
static double nu_wall(const double *T, const double *dX, const double *dX0, long lin, int type,
                      double vx, double vy, int XF, long D, double kSoot, double kAir, double *dT) {
  double Tn = T[lin];
  long nb = (type & 1) ? 1 : -1;
  long dir = ((type & 1) ? 0 : D) + lin;
  double dx = dX[dir], dx0 = dX0[dir];
  double gx = kSoot * (Tn - T[lin + nb]) / (kAir * (dx0 - dx) + kSoot * dx);
  nb = (type & 2) ? XF : -XF;
  dir = ((type & 2) ? 2 * D : 3 * D) + lin;
  dx = dX[dir];
  dx0 = dX0[dir];
  double gy = kSoot * (Tn - T[lin + nb]) / (kAir * (dx0 - dx) + kSoot * dx);
  double ax = fabs(vx), ay = fabs(vy), al = sqrt(ax * ax + ay * ay);
  *dT = gx * (ax / al) + gy * (ay / al);
  return Tn - *dT * sqrt(vx * vx + vy * vy);
}

/* The Nusselt record.  MIN_BL_BOT / MAX_BL_BOT from the walls the geometry
 * was built from, C0 (clVariablesTR.cpp:497-530); the search (in the
 * tracer's link order, MIN_BL_TOP = MIN_BL_BOT + nBL / 2) over the walls
 * now, C, and the node types now; updateNuCoeff2's type bits
 * (tfdKernels.cl:832-856); FD_calc_Nu (:919-982) with getTbm (:858-876,
 * rows x + j * nY, j < height); getNuMean's OpenCL 1.2 form (:984-1033) at
 * wavyPeriods + 1 period boundaries, its halving tree over
 * roundup(nY / 2, 32) partial sums.  Outputs as kofm_nusselt.  Returns -1
 * when no link pair brackets the domain. */
int orc_nusselt(int nX, int XF, int nY, const int16_t *nType, const double *C0, const double *C,
                const int32_t *BL, int nBL, double xMinVal, double xStopPos, int height,
                int wavyStart, int wavyPeriod, int wavyPeriods, double kSoot, double kAir,
                const double *T, const double *ux, const double *uy, const double *dX,
                const double *dX0, double *nu, double *mean, int32_t *nodes, double *vec) {
  (void)nX;
  long D = (long)XF * nY;
  int xStart = (int)floor(xMinVal), xStop = (int)floor(xStopPos);
  int n = xStop - xStart, half = nBL / 2;
  int lo = -1, hi = -1;
  for (int i = 0; i < half; i++) {
    double mid = (int)((C0[2 * BL[2 * i]] + C0[2 * BL[2 * i + 1]]) / 2.);
    if (mid >= xStart - 1) {
      lo = i;
      break;
    }
  }
  for (int i = lo + 1; lo >= 0 && i < half; i++) {
    int mid = (int)((C0[2 * BL[2 * i]] + C0[2 * BL[2 * i + 1]]) / 2.);
    if (mid > xStop + 1) {
      hi = i;
      break;
    }
  }
  if (lo < 0 || hi < 0 || n < 1) return -1;
  int perSide = hi - lo;
  int nl = (perSide + WORKGROUPSIZE_NU - 1) / WORKGROUPSIZE_NU * WORKGROUPSIZE_NU;
  if (nl > n) nl = n;
  double *best = (double *)malloc(sizeof(double) * 2 * n);
  for (int s = 0; s < 2 * n; s++) {
    best[s] = 100.;
    nodes[s] = -1;
    vec[2 * s] = vec[2 * s + 1] = 0.;
  }
  /* the tracer's link t (vtr.BL) is the wall link t on the bottom half and
   * nBL - 1 - t + nBL / 2 on the top half (iniBLinksTop,
   * clVariablesTR.cpp:600-603) */
  for (int i = 0; i < nl; i++) {
    int t = lo + i;
    if (t < 2 * half)
      nu_nearest(nY, XF, nType, C, BL, t < half ? t : nBL - 1 - t + half, xStart, xStop, xStart,
                 best, vec, nodes);
    t = half + lo + i;
    if (t < 2 * half)
      nu_nearest(nY, XF, nType, C, BL, t < half ? t : nBL - 1 - t + half, xStart, xStop,
                 xStart - n, best, vec, nodes);
  }
  free(best);
  for (int s = 0; s < 2 * n; s++) {
    int type = 0;
    if (nodes[s] != -1) {
      if (vec[2 * s] < 0.) type |= 1;
      if (vec[2 * s + 1] < 0.) type |= 2;
    }
    nodes[s] = (int32_t)((uint32_t)nodes[s] << 2) | type;
  }
  for (int ix = 0; ix < n; ix++) {
    int ii = ix + xStart;
    double Um = 0., Tm = 0.;
    for (int j = 0; j < height; j++) {
      long l = ii + (long)j * nY;
      double U = sqrt(ux[l] * ux[l] + uy[l] * uy[l]);
      Um += U;
      Tm += U * T[l];
    }
    nu[5 * ix] = Tm / Um;
    for (int w = 0; w < 2; w++) {
      int s = ix + w * n;
      int jj = nodes[s] >> 2;
      double dT = 0., Ts = 0.;
      if (jj != -1)
        Ts = nu_wall(T, dX, dX0, ii + (long)XF * jj, nodes[s] & 3, vec[2 * s], vec[2 * s + 1], XF,
                     D, kSoot, kAir, &dT);
      nu[5 * ix + 1 + 2 * w] = jj > -1 ? dT : NU_NONE;
      nu[5 * ix + 2 + 2 * w] = jj > -1 ? Ts : NU_NONE;
    }
  }
  int L = (nY / 2 + 31) / 32 * 32;
  double *ut = (double *)malloc(sizeof(double) * L), *us = (double *)malloc(sizeof(double) * L);
  for (int g = 0; g <= wavyPeriods; g++) {
    int xloc = wavyStart + g * wavyPeriod;
    for (int t = 0; t < L; t++) {
      double U1 = 0., U2 = 0., T1 = 0., T2 = 0.;
      if (2 * t < nY) {
        long l = xloc + (long)(2 * t) * XF;
        U1 = sqrt(ux[l] * ux[l] + uy[l] * uy[l]);
        T1 = T[l];
      }
      if (2 * t + 1 < nY) {
        long l = xloc + (long)(2 * t + 1) * XF;
        U2 = sqrt(ux[l] * ux[l] + uy[l] * uy[l]);
        T2 = T[l];
      }
      ut[t] = U1 * T1 + U2 * T2;
      us[t] = U1 + U2;
    }
    for (int s = L >> 1; s > 0; s >>= 1)
      for (int t = 0; t < s; t++) {
        ut[t] += ut[t + s];
        us[t] += us[t + s];
      }
    mean[2 * g] = ut[0];
    mean[2 * g + 1] = us[0];
  }
  free(ut);
  free(us);
  return 0;
}

/* ------------------------------------------------------------------------
 * Particle re-entrainment by the wall shear (TR_shear_removal,
 * trKernelsShear.cl:13-121).
 * ---------------------------------------------------------------------- */

/* The tracer's link arrays for walls C (clean walls C0):
 * updateBoundaryNodes (lsKernels.cl:246-258) on the tracer's links, whose
 * link t is the wall link t below nBL / 2 and nBL - 1 - t + nBL / 2 above
 * (iniBLinksTop, clVariablesTR.cpp:600-603). */
void orc_removal_links(int nBL, const double *C, const double *C0, const int32_t *BL, double sw,
                       double *p0, double *vN, double *len, int16_t *surf) {
  int half = nBL / 2;
  for (int t = 0; t < nBL; t++) {
    int i = t < half ? t : nBL - 1 - t + half;
    int a = BL[2 * i], b = BL[2 * i + 1];
    double vTx = C[2 * b] - C[2 * a], vTy = C[2 * b + 1] - C[2 * a + 1];
    double vNx = C0[2 * b] - C0[2 * a], vNy = C0[2 * b + 1] - C0[2 * a + 1];
    double cdx = (C[2 * a] + vTx / 2.) - (C0[2 * a] + vNx / 2.);
    double cdy = (C[2 * a + 1] + vTy / 2.) - (C0[2 * a + 1] + vNy / 2.);
    surf[t] = sqrt(cdx * cdx + cdy * cdy) >= sw ? 1 : 0;
    double l = sqrt(vTx * vTx + vTy * vTy);
    len[t] = l;
    vTx /= l;
    vTy /= l;
    vN[2 * t] = -vTy;
    vN[2 * t + 1] = vTx;
    p0[2 * t] = C[2 * a];
    p0[2 * t + 1] = C[2 * a + 1];
  }
}

/* TR_shear_removal over particles [offset, offset + count), tau per tracer
 * link (0 where TR_shear_2 writes none).  A particle whose link or size
 * index lies outside the arrays is skipped. */
void orc_shear_removal(int nSizes, const double *tauCrit, const double *dcoeff,
                       const double *lcoeff, const double *mp, double dt, double xMinVal,
                       double xMaxVal, int xIndStart, int locStride, int timerStart, int nBL,
                       const double *p0, const double *vN, const double *len, const int16_t *surf,
                       const double *tau, double *pos, const uint32_t *numRep,
                       const int16_t *type, int32_t *depFlag, uint16_t *depTimer, int32_t *loc,
                       int32_t *timer, uint32_t *blDep, long offset, long count) {
  for (long k = 0; k < count; k++) {
    long i = k + offset;
    int bl = depFlag[i];
    if (bl < 0) continue;
    int sz = type[i];
    if (bl >= nBL || sz < 0 || sz >= nSizes) continue;
    uint16_t dtim = (uint16_t)(depTimer[i] - 1);
    if (dtim == 0) {
      blDep[bl + nSizes * sz] += numRep[i];
      depFlag[i] = -2;
      loc[i] = -2;
      continue;
    }
    double shear = tau[bl];
    if (fabs(shear) <= tauCrit[sz * 2 + surf[bl]]) {
      depTimer[i] = dtim;
      continue;
    }
    double fd = dcoeff[sz], fl = lcoeff[sz] * sqrt(shear);
    double nx = vN[2 * bl], ny = vN[2 * bl + 1];
    double dx = fd * ny, dy = fd * -nx;
    double h = len[bl] * 0.5;
    double cx = p0[2 * bl] + ny * h + nx * 0.001;
    double cy = p0[2 * bl + 1] + -nx * h + ny * 0.001;
    double rx = nx * fl + dx, ry = ny * fl + dy;
    cx += rx * shear / mp[sz] * dt * dt / 2.;
    cy += ry * shear / mp[sz] * dt * dt / 2.;
    int ix = cx == cx ? (int)cx : 0, iy = cy == cy ? (int)cy : 0;
    int l = (ix - xIndStart) + iy * locStride;
    if (cx < xMinVal || cx >= xMaxVal) l = -2;
    depFlag[i] = -1;
    loc[i] = l;
    pos[2 * i] = cx;
    pos[2 * i + 1] = cy;
    timer[i] = timerStart;
  }
}
