// komega_local.hpp -- the per-node arithmetic of the k-omega part of the step,
// written once for a scalar type T: double in the forward kernels, Dual (a
// forward-mode derivative pair) for derivatives.  Evaluating a map once per
// input with that input's derivative seeded to 1 gives its transposed
// Jacobian exactly as the forward arithmetic branches (clamps, max/min, the
// difference-formula fallback).  The adjoint kernels use hand-written reverse
// passes of the same maps (koRowCoreAdj, koDiffCoreAdj, nutAdj, strainAdj;
// one pass instead of one per input), which the host checks compare with the
// forward-mode transposes, and the forward mode for the wall-spacing inputs.
//
// The functions follow the reference kernels:
//   fdCoeffsT  Update_kOmega_Diffusivities / _Coeffs_Implicit difference
//              weights (kOmegaKernels.cl:37-55, 235-257)
//   koDiffT    Update_kOmega_Diffusivities (kOmegaKernels.cl:7-95)
//   koRowT     Update_kOmega_Coeffs_Implicit (kOmegaKernels.cl:187-480)
//   nutT       eddy viscosity of LB_collision_SRT_Fluid_w_kOmega
//              (lbKernels.cl:400-430)
//   strainT    strain rate of the same kernel (lbKernels.cl:518-583)
// The header has no HIP dependency so that host code can evaluate it too.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define KOFM_HD __host__ __device__ inline
#define KOFM_ROLLED _Pragma("unroll 1")
#else
#define KOFM_HD inline
#define KOFM_ROLLED
#endif

namespace kofm {
namespace ko {

// value and derivative along one input direction
struct Dual {
  double v, d;
  KOFM_HD Dual() : v(0.), d(0.) {}
  KOFM_HD Dual(double x) : v(x), d(0.) {}
  KOFM_HD Dual(double x, double dx) : v(x), d(dx) {}
};

KOFM_HD Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
KOFM_HD Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
KOFM_HD Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
KOFM_HD Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
KOFM_HD Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
KOFM_HD Dual& operator*=(Dual& a, Dual b) { return a = a * b; }
// with a constant operand
KOFM_HD Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
KOFM_HD Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
KOFM_HD Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
KOFM_HD Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
KOFM_HD Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
KOFM_HD Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
KOFM_HD Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
KOFM_HD bool operator<(Dual a, Dual b) { return a.v < b.v; }
KOFM_HD bool operator>(Dual a, Dual b) { return a.v > b.v; }
KOFM_HD bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
KOFM_HD bool operator>=(Dual a, Dual b) { return a.v >= b.v; }

KOFM_HD double kSqrt(double x) { return sqrt(x); }
KOFM_HD double kTanh(double x) { return tanh(x); }
KOFM_HD double kAbs(double x) { return fabs(x); }
// sqrt has no derivative at 0; every use here is either multiplied by zero
// there or loses a max/min against another branch, so 0 is taken
KOFM_HD Dual kSqrt(Dual a) {
  const double s = sqrt(a.v);
  return Dual(s, s > 0. ? a.d / (2. * s) : 0.);
}
// d tanh = 1 / cosh^2 (1 - tanh^2 cancels where tanh saturates)
KOFM_HD Dual kTanh(Dual a) {
  const double c = cosh(a.v);
  return Dual(tanh(a.v), a.d / (c * c));
}
KOFM_HD Dual kAbs(Dual a) { return a.v < 0. ? -a : a; }
#if !defined(__HIPCC__)
// extended precision for host-side checks of the derivatives
inline long double kSqrt(long double x) { return sqrtl(x); }
inline long double kTanh(long double x) { return tanhl(x); }
inline long double kAbs(long double x) { return fabsl(x); }
#endif

KOFM_HD constexpr int lcx(int i) {
  return (i == 1 || i == 5 || i == 7) ? 1 : ((i == 2 || i == 6 || i == 8) ? -1 : 0);
}
KOFM_HD constexpr int lcy(int i) {
  return (i == 3 || i == 5 || i == 8) ? 1 : ((i == 4 || i == 6 || i == 7) ? -1 : 0);
}
KOFM_HD constexpr double lw(int i) { return i == 0 ? 4. / 9. : (i < 5 ? 1. / 9. : 1. / 36.); }
KOFM_HD constexpr double lw1(int c) { return c == 0 ? 2. / 3. : 1. / 6.; }

template <class T>
KOFM_HD T dphiT(int c, T u) {
  return c == 0 ? -u * u : T(0.5) * (T(double(c)) * u + u * u);
}

// equilibrium deviation feq_i - w_i in product form (the same operation
// order as the collision's geq)
template <class T>
KOFM_HD T geqT(int i, T drho, T rho, T ux, T uy) {
  const int a = lcx(i), b = lcy(i);
  const T px = dphiT(a, ux), py = dphiT(b, uy);
  return drho * T(lw(i)) + rho * (T(lw1(a)) * py + T(lw1(b)) * px + px * py);
}

// Three-point weights for the spacings dp (forward) and dm (backward):
// c[0..2] first derivative (next, previous, centre), c[3..5] second.
template <class T>
KOFM_HD void fdCoeffsT(T dp, T dm, T (&c)[6]) {
  dp = dp > T(0.1) ? dp : T(0.1);
  dm = dm > T(0.1) ? dm : T(0.1);
  T dc = dp + dm;
  dc = dc > T(0.1) ? dc : T(0.1);
  c[0] = dm / (dp * dc);
  c[1] = -dp / (dm * dc);
  c[2] = (dp - dm) / (dp * dm);
  if (kAbs(c[2]) > T(10)) {
    c[0] = T(1) / dc;
    c[1] = T(-1) / dc;
    c[2] = T(0);
  }
  c[3] = T(2) / (dp * dc);
  c[4] = T(2) / (dm * dc);
  c[5] = T(-2) / (dp * dm);
}

// Stencil order of every five-point input: node, E, W, N, S.
// Inputs of the diffusivity map at one node (k and omega after the
// collision's clamps).
enum DiffInput {
  kDiffK = 0,     // 5: k at the stencil
  kDiffO = 5,     // 5: omega at the stencil
  kDiffNut = 10,  // eddy viscosity
  kDiffYw = 11,   // wall distance
  kDiffRho = 12,  // density
  kDiffDx = 13,   // 4: wall spacings E, W, N, S
  kDiffInputs = 17
};
enum DiffOutput { kDiffMx = 0, kDiffMy, kDiffF1, kDiffDk, kDiffDo, kDiffOutputs };

// The map with its difference weights given (C = double holds them
// constant; the spacing inputs are then not read).
template <class T, class C>
KOFM_HD void koDiffCore(const T (&in)[kDiffInputs], const C (&ax)[6], const C (&ay)[6], double nu,
                        T (&o)[kDiffOutputs]) {
  const T* K = in + kDiffK;
  const T* O = in + kDiffO;
  const T om = O[0], k = K[0], yw = in[kDiffYw];
  const T dOdx = ax[0] * O[1] + ax[1] * O[2] + ax[2] * om;
  const T dKdx = ax[0] * K[1] + ax[1] * K[2] + ax[2] * k;
  const T dOdy = ay[0] * O[3] + ay[1] * O[4] + ay[2] * om;
  const T dKdy = ay[0] * K[3] + ay[1] * K[4] + ay[2] * k;
  const T mlt = T(2 * 0.856) / om;
  const T mx = mlt * dKdx, my = mlt * dKdy;
  o[kDiffMx] = mx;
  o[kDiffMy] = my;
  T cd = in[kDiffRho] * (mx * dOdx + my * dOdy);
  cd = cd > T(1.0e-10) ? cd : T(1.0e-10);
  const T a = kSqrt(k) / (T(0.09) * om * yw), b = T(500) * T(nu) / (yw * yw * om);
  T arg = a > b ? a : b;
  const T c = T(4 * 0.856) * k / (cd * yw * yw);
  arg = arg < c ? arg : c;
  const T a2 = arg * arg;
  const T F1 = kTanh(a2 * a2);
  o[kDiffF1] = F1;
  const T nt = in[kDiffNut];
  o[kDiffDk] = T(nu) + (T(0.85) * F1 + T(1) * (T(1) - F1)) * nt;
  o[kDiffDo] = T(nu) + (T(0.5) * F1 + T(0.856) * (T(1) - F1)) * nt;
}

template <class T>
KOFM_HD void koDiffT(const T (&in)[kDiffInputs], double nu, T (&o)[kDiffOutputs]) {
  T ax[6], ay[6];
  fdCoeffsT<T>(in[kDiffDx], in[kDiffDx + 1], ax);
  fdCoeffsT<T>(in[kDiffDx + 2], in[kDiffDx + 3], ay);
  koDiffCore<T, T>(in, ax, ay, nu, o);
}

// Inputs of one node's pair of normalised rows.
enum RowInput {
  kRowDk = 0,    // 5: k diffusivity at the stencil
  kRowDo = 5,    // 5: omega diffusivity at the stencil
  kRowMx = 10,   // cross-diffusion m grad k, x and y
  kRowMy = 11,
  kRowUx = 12,   // stored velocity (u + F/2), density
  kRowUy = 13,
  kRowRho = 14,
  kRowK = 15,    // k, omega after the collision's clamps
  kRowO = 16,
  kRowF1 = 17,
  kRowNut = 18,
  kRowS = 19,    // 3: strain Sxx, Sxy, Syy of this step
  kRowYw = 22,
  kRowDx = 23,   // 4: wall spacings
  kRowInputs = 27
};
// Outputs: E, W, N, S coefficients of the k row, of the omega row, then the
// two right-hand sides.
enum RowOutput { kRowAk = 0, kRowAo = 4, kRowBk = 8, kRowBo = 9, kRowOutputs = 10 };

template <class T, class C>
KOFM_HD void koRowCore(const T (&in)[kRowInputs], const C (&ax)[6], const C (&ay)[6], bool nesw,
                       double nu, double dtIn, T (&o)[kRowOutputs]) {
  const T* Dk = in + kRowDk;
  const T* Do = in + kRowDo;
  const T dk = Dk[0], dw = Do[0];
  const T dkdx = ax[0] * Dk[1] + ax[1] * Dk[2] + ax[2] * dk;
  const T dwdx = ax[0] * Do[1] + ax[1] * Do[2] + ax[2] * dw;
  const T dkdy = ay[0] * Dk[3] + ay[1] * Dk[4] + ay[2] * dk;
  const T dwdy = ay[0] * Do[3] + ay[1] * Do[4] + ay[2] * dw;
  const T Ux = in[kRowUx], Uy = in[kRowUy], Rh = in[kRowRho];
  const T Jxw = Ux - dwdx - in[kRowMx], Jyw = Uy - dwdy - in[kRowMy];
  const T Jxk = Ux - dkdx, Jyk = Uy - dkdy;
  T om = in[kRowO] > T(0) ? in[kRowO] : T(0);
  const T k = in[kRowK] > T(0) ? in[kRowK] : T(0);
  const T F1 = in[kRowF1], nt = in[kRowNut];
  const T sxx = in[kRowS], sxy = in[kRowS + 1], syy = in[kRowS + 2];
  T Pk = T(2) * nt * (sxx * sxx + syy * syy + T(2) * sxy * sxy) -
         T(2. / 3.) * k * Rh * (sxx + syy);
  const T lim = T(10 * 0.09) * k * om;
  Pk = Pk < lim ? Pk : lim;
  Pk = Pk > T(0) ? Pk : T(0);
  const T gam = (F1 * T(5. / 9.) + (T(1) - F1) * T(0.44)) / nt;
  T Sc = T(-0.09) * om;
  T dt = T(dtIn);
  const T Kc = T(1) + dt * (Jxk * ax[2] + Jyk * ay[2] - dk * (ax[5] + ay[5])) - Sc;
  T iv = T(1) / Kc;
  o[kRowAk + 0] = dt * (Jxk * ax[0] - dk * ax[3]) * iv;
  o[kRowAk + 1] = dt * (Jxk * ax[1] - dk * ax[4]) * iv;
  o[kRowAk + 2] = dt * (Jyk * ay[0] - dk * ay[3]) * iv;
  o[kRowAk + 3] = dt * (Jyk * ay[1] - dk * ay[4]) * iv;
  o[kRowBk] = (k + dt * Pk) * iv;
  if (nesw) {  // omega wall row (kOmegaKernels.cl:389-403)
    dt = T(0);
    Sc = T(0);
    const T yp = in[kRowYw];
    const T ov = (T(6) * T(nu) / T(0.075)) / (yp * yp);
    const T ol = kSqrt(k) / (T(0.5477225575 * 0.41) * yp);
    const T ow = kSqrt(ov * ov + ol * ol);
    om = ow < T(1000) ? ow : T(1000);
  }
  const T Wc = T(1) + dt * (Jxw * ax[2] + Jyw * ay[2] - dw * (ax[5] + ay[5])) - Sc;
  iv = T(1) / Wc;
  o[kRowAo + 0] = dt * (Jxw * ax[0] - dw * ax[3]) * iv;
  o[kRowAo + 1] = dt * (Jxw * ax[1] - dw * ax[4]) * iv;
  o[kRowAo + 2] = dt * (Jyw * ay[0] - dw * ay[3]) * iv;
  o[kRowAo + 3] = dt * (Jyw * ay[1] - dw * ay[4]) * iv;
  o[kRowBo] = (om + dt * (gam * Pk)) * iv;
}

template <class T>
KOFM_HD void koRowT(const T (&in)[kRowInputs], bool nesw, double nu, double dtIn,
                    T (&o)[kRowOutputs]) {
  T ax[6], ay[6];
  fdCoeffsT<T>(in[kRowDx], in[kRowDx + 1], ax);
  fdCoeffsT<T>(in[kRowDx + 2], in[kRowDx + 3], ay);
  koRowCore<T, T>(in, ax, ay, nesw, nu, dtIn, o);
}

// Reverse mode of koRowCore with constant difference weights: xb[i] =
// (J^T yb)_i for every input except the spacings (kRowDx..), in one pass.
// Branches follow the forward's comparisons; d sqrt(k) at k = 0 is taken as
// 0 and d tanh as 1 / cosh^2, as in the dual numbers.
KOFM_HD void koRowCoreAdj(const double (&in)[kRowInputs], const double (&ax)[6],
                          const double (&ay)[6], bool nesw, double nu, double dtIn,
                          const double (&yb)[kRowOutputs], double (&xb)[kRowInputs]) {
  const double* Dk = in + kRowDk;
  const double* Do = in + kRowDo;
  const double dk = Dk[0], dw = Do[0];
  const double dkdx = ax[0] * Dk[1] + ax[1] * Dk[2] + ax[2] * dk;
  const double dwdx = ax[0] * Do[1] + ax[1] * Do[2] + ax[2] * dw;
  const double dkdy = ay[0] * Dk[3] + ay[1] * Dk[4] + ay[2] * dk;
  const double dwdy = ay[0] * Do[3] + ay[1] * Do[4] + ay[2] * dw;
  const double Ux = in[kRowUx], Uy = in[kRowUy], Rh = in[kRowRho];
  const double Jxw = Ux - dwdx - in[kRowMx], Jyw = Uy - dwdy - in[kRowMy];
  const double Jxk = Ux - dkdx, Jyk = Uy - dkdy;
  const double om0 = in[kRowO] > 0. ? in[kRowO] : 0.;
  const double k = in[kRowK] > 0. ? in[kRowK] : 0.;
  const double F1 = in[kRowF1], nt = in[kRowNut];
  const double sxx = in[kRowS], sxy = in[kRowS + 1], syy = in[kRowS + 2];
  const double Q = sxx * sxx + syy * syy + 2. * sxy * sxy, Tr = sxx + syy;
  const double PkRaw = 2. * nt * Q - (2. / 3.) * k * Rh * Tr;
  const double lim = 0.9 * k * om0;
  const bool rawWins = PkRaw < lim;
  const double Pk1 = rawWins ? PkRaw : lim;
  const bool pos = Pk1 > 0.;
  const double Pk = pos ? Pk1 : 0.;
  const double gam = (F1 * (5. / 9.) + (1. - F1) * 0.44) / nt;
  const double Sc = -0.09 * om0;
  const double dt = dtIn;
  const double Kc = 1. + dt * (Jxk * ax[2] + Jyk * ay[2] - dk * (ax[5] + ay[5])) - Sc;
  const double ivk = 1. / Kc;
  const double Nk[4] = {dt * (Jxk * ax[0] - dk * ax[3]), dt * (Jxk * ax[1] - dk * ax[4]),
                        dt * (Jyk * ay[0] - dk * ay[3]), dt * (Jyk * ay[1] - dk * ay[4])};
  double dtw = dt, om = om0, ov = 0., ol = 0., ow = 0.;
  const double yp = in[kRowYw];
  if (nesw) {
    dtw = 0.;
    ov = (6. * nu / 0.075) / (yp * yp);
    ol = sqrt(k) / ((0.5477225575 * 0.41) * yp);
    ow = sqrt(ov * ov + ol * ol);
    om = ow < 1000. ? ow : 1000.;
  }
  const double Scw = nesw ? 0. : Sc;
  const double Wc = 1. + dtw * (Jxw * ax[2] + Jyw * ay[2] - dw * (ax[5] + ay[5])) - Scw;
  const double ivw = 1. / Wc;
  const double Nw[4] = {dtw * (Jxw * ax[0] - dw * ax[3]), dtw * (Jxw * ax[1] - dw * ax[4]),
                        dtw * (Jyw * ay[0] - dw * ay[3]), dtw * (Jyw * ay[1] - dw * ay[4])};
  // k row
  const double* Akb = yb + kRowAk;
  const double* Aob = yb + kRowAo;
  const double bkb = yb[kRowBk], bob = yb[kRowBo];
  double ivkb = bkb * (k + dt * Pk);
  for (int d = 0; d < 4; d++) ivkb += Akb[d] * Nk[d];
  double Jxkb = dt * ivk * (Akb[0] * ax[0] + Akb[1] * ax[1]);
  double Jykb = dt * ivk * (Akb[2] * ay[0] + Akb[3] * ay[1]);
  double dkb = -dt * ivk * (Akb[0] * ax[3] + Akb[1] * ax[4] + Akb[2] * ay[3] + Akb[3] * ay[4]);
  double kb = bkb * ivk;
  double Pkb = bkb * dt * ivk;
  const double Kcb = -ivkb * ivk * ivk;
  Jxkb += Kcb * dt * ax[2];
  Jykb += Kcb * dt * ay[2];
  dkb -= Kcb * dt * (ax[5] + ay[5]);
  double Scb = -Kcb;
  // omega row
  double ivwb = bob * (om + dtw * (gam * Pk));
  for (int d = 0; d < 4; d++) ivwb += Aob[d] * Nw[d];
  double Jxwb = dtw * ivw * (Aob[0] * ax[0] + Aob[1] * ax[1]);
  double Jywb = dtw * ivw * (Aob[2] * ay[0] + Aob[3] * ay[1]);
  double dwb = -dtw * ivw * (Aob[0] * ax[3] + Aob[1] * ax[4] + Aob[2] * ay[3] + Aob[3] * ay[4]);
  const double omb = bob * ivw;
  const double gamb = bob * ivw * dtw * Pk;
  Pkb += bob * ivw * dtw * gam;
  const double Wcb = -ivwb * ivw * ivw;
  Jxwb += Wcb * dtw * ax[2];
  Jywb += Wcb * dtw * ay[2];
  dwb -= Wcb * dtw * (ax[5] + ay[5]);
  double om0b = 0., ypb = 0.;
  if (nesw) {
    if (ow < 1000. && ow > 0.) {
      const double ovb = omb * ov / ow, olb = omb * ol / ow;
      ypb += ovb * (-2. * ov / yp) + olb * (-ol / yp);
      const double sk = sqrt(k);
      if (sk > 0.) kb += olb / (2. * sk * (0.5477225575 * 0.41) * yp);
    }
  } else {
    Scb += -Wcb;
    om0b += omb;
  }
  om0b += -0.09 * Scb;
  double F1b = gamb * (5. / 9. - 0.44) / nt;
  double ntb = -gamb * gam / nt;
  const double Pk1b = pos ? Pkb : 0.;
  const double PkRawb = rawWins ? Pk1b : 0.;
  const double limb = rawWins ? 0. : Pk1b;
  kb += limb * 0.9 * om0;
  om0b += limb * 0.9 * k;
  ntb += PkRawb * 2. * Q;
  kb += PkRawb * (-(2. / 3.) * Rh * Tr);
  const double Rhb = PkRawb * (-(2. / 3.) * k * Tr);
  const double c23 = (2. / 3.) * k * Rh;
  const double sxxb = PkRawb * (4. * nt * sxx - c23);
  const double syyb = PkRawb * (4. * nt * syy - c23);
  const double sxyb = PkRawb * (8. * nt * sxy);
  const double dkdxb = -Jxkb, dkdyb = -Jykb, dwdxb = -Jxwb, dwdyb = -Jywb;
  xb[kRowDk + 0] = dkb + ax[2] * dkdxb + ay[2] * dkdyb;
  xb[kRowDk + 1] = ax[0] * dkdxb;
  xb[kRowDk + 2] = ax[1] * dkdxb;
  xb[kRowDk + 3] = ay[0] * dkdyb;
  xb[kRowDk + 4] = ay[1] * dkdyb;
  xb[kRowDo + 0] = dwb + ax[2] * dwdxb + ay[2] * dwdyb;
  xb[kRowDo + 1] = ax[0] * dwdxb;
  xb[kRowDo + 2] = ax[1] * dwdxb;
  xb[kRowDo + 3] = ay[0] * dwdyb;
  xb[kRowDo + 4] = ay[1] * dwdyb;
  xb[kRowMx] = -Jxwb;
  xb[kRowMy] = -Jywb;
  xb[kRowUx] = Jxkb + Jxwb;
  xb[kRowUy] = Jykb + Jywb;
  xb[kRowRho] = Rhb;
  xb[kRowK] = in[kRowK] > 0. ? kb : 0.;
  xb[kRowO] = in[kRowO] > 0. ? om0b : 0.;
  xb[kRowF1] = F1b;
  xb[kRowNut] = ntb;
  xb[kRowS] = sxxb;
  xb[kRowS + 1] = sxyb;
  xb[kRowS + 2] = syyb;
  xb[kRowYw] = ypb;
}

// Reverse mode of koDiffCore with constant difference weights: xb[i] for
// every input except the spacings (kDiffDx..), in one pass.
KOFM_HD void koDiffCoreAdj(const double (&in)[kDiffInputs], const double (&ax)[6],
                           const double (&ay)[6], double nu, const double (&yb)[kDiffOutputs],
                           double (&xb)[kDiffInputs]) {
  const double* K = in + kDiffK;
  const double* O = in + kDiffO;
  const double om = O[0], k = K[0], yw = in[kDiffYw], rho = in[kDiffRho], nt = in[kDiffNut];
  const double dOdx = ax[0] * O[1] + ax[1] * O[2] + ax[2] * om;
  const double dKdx = ax[0] * K[1] + ax[1] * K[2] + ax[2] * k;
  const double dOdy = ay[0] * O[3] + ay[1] * O[4] + ay[2] * om;
  const double dKdy = ay[0] * K[3] + ay[1] * K[4] + ay[2] * k;
  const double mlt = (2 * 0.856) / om;
  const double mx = mlt * dKdx, my = mlt * dKdy;
  const double cdRaw = rho * (mx * dOdx + my * dOdy);
  const bool cdFree = cdRaw > 1.0e-10;
  const double cd = cdFree ? cdRaw : 1.0e-10;
  const double sk = sqrt(k);
  const double a = sk / (0.09 * om * yw), b = 500. * nu / (yw * yw * om);
  const bool aWins = a > b;
  const double arg1 = aWins ? a : b;
  const double c = (4 * 0.856) * k / (cd * yw * yw);
  const bool argFree = arg1 < c;
  const double arg = argFree ? arg1 : c;
  const double a2 = arg * arg, a4 = a2 * a2;
  const double F1 = tanh(a4);
  const double F1b = yb[kDiffF1] + yb[kDiffDk] * (0.85 - 1.) * nt + yb[kDiffDo] * (0.5 - 0.856) * nt;
  const double ntb = yb[kDiffDk] * (0.85 * F1 + (1. - F1)) + yb[kDiffDo] * (0.5 * F1 + 0.856 * (1. - F1));
  const double ch = cosh(a4);
  const double argb = F1b / (ch * ch) * 2. * a2 * 2. * arg;
  const double arg1b = argFree ? argb : 0.;
  const double cb = argFree ? 0. : argb;
  double kb = cb * (4 * 0.856) / (cd * yw * yw);
  double ywb = cb * (-2. * c / yw);
  const double cdb = cb * (-c / cd);
  double omb = 0.;
  if (aWins) {
    if (sk > 0.) kb += arg1b * a / (2. * k);
    omb += arg1b * (-a / om);
    ywb += arg1b * (-a / yw);
  } else {
    ywb += arg1b * (-2. * b / yw);
    omb += arg1b * (-b / om);
  }
  const double cdRawb = cdFree ? cdb : 0.;
  const double rhob = cdRawb * (mx * dOdx + my * dOdy);
  const double mxb = yb[kDiffMx] + cdRawb * rho * dOdx;
  const double myb = yb[kDiffMy] + cdRawb * rho * dOdy;
  const double dOdxb = cdRawb * rho * mx, dOdyb = cdRawb * rho * my;
  const double mltb = mxb * dKdx + myb * dKdy;
  const double dKdxb = mxb * mlt, dKdyb = myb * mlt;
  omb += mltb * (-mlt / om);
  xb[kDiffK + 0] = kb + ax[2] * dKdxb + ay[2] * dKdyb;
  xb[kDiffK + 1] = ax[0] * dKdxb;
  xb[kDiffK + 2] = ax[1] * dKdxb;
  xb[kDiffK + 3] = ay[0] * dKdyb;
  xb[kDiffK + 4] = ay[1] * dKdyb;
  xb[kDiffO + 0] = omb + ax[2] * dOdxb + ay[2] * dOdyb;
  xb[kDiffO + 1] = ax[0] * dOdxb;
  xb[kDiffO + 2] = ax[1] * dOdxb;
  xb[kDiffO + 3] = ay[0] * dOdyb;
  xb[kDiffO + 4] = ay[1] * dOdyb;
  xb[kDiffNut] = ntb;
  xb[kDiffYw] = ywb;
  xb[kDiffRho] = rhob;
}

// Eddy viscosity from k and omega after the clamps, the wall distance and
// the strain of the previous step.
enum NutInput { kNutK = 0, kNutO, kNutYw, kNutS, kNutInputs = 6 };

template <class T>
KOFM_HD T nutT(const T (&in)[kNutInputs], double nu, double maxNut) {
  const T k = in[kNutK], om = in[kNutO], yw = in[kNutYw];
  const T sxx = in[kNutS], sxy = in[kNutS + 1], syy = in[kNutS + 2];
  const T a = T(2) * kSqrt(k) / (T(0.09) * om * yw);
  const T b = T(500) * T(nu) / (yw * yw * om);
  const T arg2 = a > b ? a : b;
  T F2 = kTanh(arg2 * arg2);
  F2 *= kSqrt(T(2) * (sxx * sxx + T(2) * sxy * sxy + syy * syy));
  const T den = T(0.31) * om > F2 ? T(0.31) * om : F2;
  T nt = T(0.31) * k / den;
  nt = nt < T(maxNut) ? nt : T(maxNut);
  nt = nt > T(1e-16) ? nt : T(1e-16);
  return nt;
}

// Strain Sxx, Sxy, Syy from the nine pre-collision deviations (FP64), with
// the laminar relaxation frequency omega0.
template <class T>
KOFM_HD void strainT(const T (&gd)[9], double omega0, T (&s)[3]) {
  const T dr = ((gd[0] + gd[1]) + (gd[2] + gd[3])) + ((gd[4] + gd[5]) + (gd[6] + gd[7])) + gd[8];
  const T rd = T(1) + dr;
  const T uxd = (((gd[1] - gd[2]) + (gd[5] - gd[6])) + (gd[7] - gd[8])) / rd;
  const T uyd = (((gd[3] - gd[4]) + (gd[5] - gd[6])) - (gd[7] - gd[8])) / rd;
  T ne[9];
  ne[0] = T(0);
  for (int i = 1; i < 9; i++) ne[i] = gd[i] - geqT<T>(i, dr, rd, uxd, uyd);
  const T c = T(-1.5) * T(omega0) / rd;
  const T exx = ((ne[1] + ne[2]) + (ne[5] + ne[6])) + (ne[7] + ne[8]);
  const T eyy = ((ne[3] + ne[4]) + (ne[5] + ne[6])) + (ne[7] + ne[8]);
  const T exy = (ne[5] + ne[6]) - (ne[7] + ne[8]);
  s[0] = c * exx;
  s[1] = c * exy;
  s[2] = c * eyy;
}

// sink(i, (J^T yb)_i) for the inputs lo <= i < hi of the local map f: NI
// inputs -> NO outputs at the point x, one forward-mode evaluation per input.
template <int NI, int NO, class F, class Sink>
KOFM_HD void transposeRange(const double (&x)[NI], const double (&yb)[NO], int lo, int hi, F&& f,
                            Sink&& sink) {
  KOFM_ROLLED
  for (int i = lo; i < hi; i++) {
    Dual in[NI];
    for (int j = 0; j < NI; j++) in[j] = Dual(x[j], j == i ? 1. : 0.);
    Dual out[NO];
    f(in, out);
    double s = 0.;
    for (int o = 0; o < NO; o++) s += yb[o] * out[o].d;
    sink(i, s);
  }
}

// Reverse mode of nutT: xb = (d nut / d in) * nb, in one pass.
KOFM_HD void nutAdj(const double (&in)[kNutInputs], double nu, double maxNut, double nb,
                    double (&xb)[kNutInputs]) {
  const double k = in[kNutK], om = in[kNutO], yw = in[kNutYw];
  const double sxx = in[kNutS], sxy = in[kNutS + 1], syy = in[kNutS + 2];
  const double sk = sqrt(k);
  const double a = 2. * sk / (0.09 * om * yw);
  const double b = 500. * nu / (yw * yw * om);
  const bool aWins = a > b;
  const double arg2 = aWins ? a : b;
  const double F2 = tanh(arg2 * arg2);
  const double Sm = sqrt(2. * (sxx * sxx + 2. * sxy * sxy + syy * syy));
  const double F2S = F2 * Sm;
  const bool omWins = 0.31 * om > F2S;
  const double den = omWins ? 0.31 * om : F2S;
  const double nt0 = 0.31 * k / den;
  const bool underMax = nt0 < maxNut;
  const double nt1 = underMax ? nt0 : maxNut;
  const double nt1b = nt1 > 1e-16 ? nb : 0.;
  const double nt0b = underMax ? nt1b : 0.;
  double kb = nt0b * 0.31 / den;
  const double denb = -nt0b * nt0 / den;
  double omb = omWins ? denb * 0.31 : 0.;
  const double F2Sb = omWins ? 0. : denb;
  const double F2b = F2Sb * Sm, Smb = F2Sb * F2;
  // Sm = sqrt(2 Q): dSm/dQ = 1 / Sm
  const double Qb = Sm > 0. ? Smb / Sm : 0.;
  const double ch = cosh(arg2 * arg2);
  const double arg2b = F2b / (ch * ch) * 2. * arg2;
  double ywb = 0.;
  if (aWins) {
    if (sk > 0.) kb += arg2b * a / (2. * k);
    omb -= arg2b * a / om;
    ywb -= arg2b * a / yw;
  } else {
    ywb -= arg2b * 2. * b / yw;
    omb -= arg2b * b / om;
  }
  xb[kNutK] = kb;
  xb[kNutO] = omb;
  xb[kNutYw] = ywb;
  xb[kNutS] = Qb * 2. * sxx;
  xb[kNutS + 1] = Qb * 4. * sxy;
  xb[kNutS + 2] = Qb * 2. * syy;
}

// Reverse mode of strainT: gb = (d S / d gd)^T sb, in one pass.
KOFM_HD void strainAdj(const double (&gd)[9], double omega0, const double (&sb)[3], double (&gb)[9]) {
  const double dr = ((gd[0] + gd[1]) + (gd[2] + gd[3])) + ((gd[4] + gd[5]) + (gd[6] + gd[7])) + gd[8];
  const double rd = 1. + dr;
  const double jx = ((gd[1] - gd[2]) + (gd[5] - gd[6])) + (gd[7] - gd[8]);
  const double jy = ((gd[3] - gd[4]) + (gd[5] - gd[6])) - (gd[7] - gd[8]);
  const double uxd = jx / rd, uyd = jy / rd;
  double ne[9];
  ne[0] = 0.;
  for (int i = 1; i < 9; i++) ne[i] = gd[i] - geqT<double>(i, dr, rd, uxd, uyd);
  const double c = -1.5 * omega0 / rd;
  const double exx = ((ne[1] + ne[2]) + (ne[5] + ne[6])) + (ne[7] + ne[8]);
  const double eyy = ((ne[3] + ne[4]) + (ne[5] + ne[6])) + (ne[7] + ne[8]);
  const double exy = (ne[5] + ne[6]) - (ne[7] + ne[8]);
  const double cb = sb[0] * exx + sb[1] * exy + sb[2] * eyy;
  const double exxb = c * sb[0], exyb = c * sb[1], eyyb = c * sb[2];
  double drb = 0., rdb = -cb * c / rd, uxb = 0., uyb = 0.;
  gb[0] = 0.;
  for (int i = 1; i < 9; i++) {
    const int ax = lcx(i), ay = lcy(i);
    const double neb = (ax != 0 ? exxb : 0.) + (ay != 0 ? eyyb : 0.) +
                       (i == 5 || i == 6 ? exyb : (i == 7 || i == 8 ? -exyb : 0.));
    gb[i] = neb;
    // geq = dr w_i + rd (w1(a) py + w1(b) px + px py)
    const double px = dphiT(ax, uxd), py = dphiT(ay, uyd);
    const double dpx = ax == 0 ? -2. * uxd : 0.5 * (double(ax) + 2. * uxd);
    const double dpy = ay == 0 ? -2. * uyd : 0.5 * (double(ay) + 2. * uyd);
    drb -= neb * lw(i);
    rdb -= neb * (lw1(ax) * py + lw1(ay) * px + px * py);
    uxb -= neb * rd * (lw1(ay) + py) * dpx;
    uyb -= neb * rd * (lw1(ax) + px) * dpy;
  }
  const double jxb = uxb / rd, jyb = uyb / rd;
  rdb -= (uxb * uxd + uyb * uyd) / rd;
  drb += rdb;
  for (int i = 0; i < 9; i++) gb[i] += drb;
  gb[1] += jxb; gb[2] -= jxb; gb[5] += jxb; gb[6] -= jxb; gb[7] += jxb; gb[8] -= jxb;
  gb[3] += jyb; gb[4] -= jyb; gb[5] += jyb; gb[6] -= jyb; gb[7] -= jyb; gb[8] += jyb;
}

// xb += J^T yb for the local map f: NI inputs -> NO outputs at the point x,
// one forward-mode evaluation per input (inputs with skip(i) true are left
// alone).
template <int NI, int NO, class F, class Skip>
KOFM_HD void transposeApply(const double (&x)[NI], const double (&yb)[NO], double (&xb)[NI], F&& f,
                            Skip&& skip) {
  // one evaluation of f per input, kept as a loop (unrolled, the NI
  // evaluations would not fit the register file)
  KOFM_ROLLED
  for (int i = 0; i < NI; i++) {
    if (skip(i)) continue;
    Dual in[NI];
    for (int j = 0; j < NI; j++) in[j] = Dual(x[j], j == i ? 1. : 0.);
    Dual out[NO];
    f(in, out);
    double s = 0.;
    for (int o = 0; o < NO; o++) s += yb[o] * out[o].d;
    xb[i] += s;
  }
}

}  // namespace ko
}  // namespace kofm
