// init.cpp -- clVariablesLB::iniDists(Umax) (clVariablesLB.cpp:427-496): the
// initial velocity field of a channel run, laminar (a parabola between the
// walls of each column) or turbulent (the law of the wall), optionally
// perturbed by streamwise streaks and wall-normal waves whose amplitudes are
// drawn from the reference's Box-Muller generator.
#include "init.hpp"

#include <cmath>
#include <cstdint>

namespace kofm {
namespace {

// rand()/srand() of the C library the reference is built with (MSVC):
// x <- 214013 x + 2531011 (mod 2^32), rand() = bits 16..30 of x,
// RAND_MAX = 32767.
struct CRand {
  uint32_t x;
  int next() {
    x = x * 214013u + 2531011u;
    return int((x >> 16) & 0x7FFFu);
  }
};

// boxMuller (HelperFuncs.cpp:107-126): the polar method, keeping the second
// deviate for the next call.
struct BoxMuller {
  CRand r;
  bool haveSpare = false;
  double spare = 0.;
  double operator()(double mean, double sd) {
    double y;
    if (haveSpare) {
      y = spare;
      haveSpare = false;
    } else {
      double a, b, s;
      do {
        a = 2. * double(r.next()) / 32767. - 1.;
        b = 2. * double(r.next()) / 32767. - 1.;
        s = a * a + b * b;
      } while (s >= 1.);
      const double m = std::sqrt(-2. * std::log(s) / s);
      y = a * m;
      spare = b * m;
      haveSpare = true;
    }
    return mean + y * sd;
  }
};

}  // namespace

void iniPoiseuille(const Geometry& g, const kofm_ini_flow& p, double* rho, double* ux, double* uy) {
  const double R = p.pipe_radius, nu = p.nu, kPi = 3.14159265358979323846;
  // friction and bulk scales of the reference (clVariablesLB.cpp:432-447)
  const double uparaMax = p.Re * nu / R * 1.5;
  const double utau = std::sqrt(p.Re * 3.) * nu / R;
  const double ubar = uparaMax * 2. / 3.;
  const double duPlus = ubar * p.duplus_mult / utau;
  const double betaPlus = 2. * kPi / p.alpha_mult;  // the reference pairs beta with Alpha
  const double alphaPlus = 2. * kPi / p.beta_mult;
  const double epsPlus = ubar / p.epsilon_mult;
  BoxMuller bm{CRand{p.seed}};
  const int64_t D = g.D;
  const int XF = g.XF;
  for (int i = 0; i < g.nX; i++) {
    int j = 0;
    while (j < g.nY && (g.nType[size_t(i) + size_t(XF) * j] & kMSolid)) j++;
    if (j >= g.nY) continue;
    // the first fluid node's distance to the wall below (dXArr S plane)
    double Y = g.dX[size_t(i) + size_t(XF) * j + size_t(3 * D)];
    for (; j < g.nY && (g.nType[size_t(i) + size_t(XF) * j] & kMFluid); j++, Y += 1.) {
      const double dev = bm(1., 0.2);
      const double ypara = Y < R ? R - Y : Y - R;
      const double yturb = Y < R ? Y : 2. * R - Y;
      const double yplus = yturb * utau / nu;
      const double xplus = double(i) * utau / nu;
      // viscous sublayer below y+ = 11.445..., where it meets the log law
      const double uplus = yplus < 11.44532166 ? yplus : std::log(yplus) / 0.41 + 5.5;
      double u = p.turbulent_profile ? uplus * utau : uparaMax * (1. - ypara * ypara / R / R);
      double v = 0.;
      if (p.perturb) {
        u += (utau * duPlus / 2.) * (yplus / 40.) * std::exp(-p.sigma_mult * yplus * yplus + 0.5) *
             std::cos(betaPlus * xplus) * dev;
        v = yplus * epsPlus * std::exp(-p.sigma_mult * yplus * yplus) * std::sin(alphaPlus * xplus) *
            dev;
      }
      const size_t n = size_t(i) + size_t(XF) * j;
      rho[n] = p.rho;
      ux[n] = u;
      uy[n] = v;
    }
  }
}

}  // namespace kofm
