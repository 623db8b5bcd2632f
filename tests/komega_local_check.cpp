// Host check of the k-omega local maps' transposed Jacobians (built and run
// by test_komega_local.py with g++): for random realistic inputs, J^T yb
// from the forward-mode evaluation must match central differences of
// <yb, f(x)> in every input direction.
#include <cmath>
#include <cstdio>
#include <random>

#include "komega_local.hpp"

using namespace kofm::ko;

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }

template <int NI, int NO, class F>
static double check(const char* name, const double (&x)[NI], F&& f) {
  double yb[NO];
  for (int o = 0; o < NO; o++) yb[o] = U(-1., 1.);
  double xb[NI] = {};
  transposeApply<NI, NO>(x, yb, xb, [&](const Dual (&in)[NI], Dual (&out)[NO]) { f(in, out); },
                         [](int) { return false; });
  double worst = 0.;
  long double fds[NI], total = 0.L;
  for (int i = 0; i < NI; i++) {
    // central difference in extended precision
    const long double h = 1e-7L * (std::fabs(x[i]) > 1e-30 ? std::fabs(x[i]) : 1.);
    long double xp[NI], xm[NI], yp[NO], ym[NO];
    for (int j = 0; j < NI; j++) xp[j] = xm[j] = x[j];
    xp[i] += h;
    xm[i] -= h;
    f(xp, yp);
    f(xm, ym);
    long double fd = 0.L;
    for (int o = 0; o < NO; o++) fd += yb[o] * (yp[o] - ym[o]) / (2.L * h);
    fds[i] = fd;
    total += fabsl(fd) * (std::fabs(x[i]) > 1e-30 ? std::fabs(x[i]) : 1.);
  }
  for (int i = 0; i < NI; i++) {
    // relative error, or (where the double forward is itself ill-conditioned,
    // e.g. through a cancelling cross-diffusion term) error small against the
    // response of <yb, f> to relative perturbations of all inputs
    const double fd = double(fds[i]), xs = std::fabs(x[i]) > 1e-30 ? std::fabs(x[i]) : 1.;
    const double rel = std::fabs(xb[i] - fd) / (std::fabs(fd) + 1e-300);
    const double scaled = std::fabs(xb[i] - fd) * xs / (double(total) + 1e-300);
    const double err = rel < 1e-6 ? 0. : scaled * 1e3;
    if (err > worst) worst = err;
    if (err > 1e-6)
      std::printf("%s input %d: transpose %.12e fd %.12e (scaled %.2e)\n", name, i, xb[i], fd, scaled);
  }
  return worst;
}

int main() {
  int bad = 0;
  for (int trial = 0; trial < 200; trial++) {
    const double nu = U(0.002, 0.02), dt = 1.;
    // diffusivities
    double dx[kDiffInputs];
    for (int j = 0; j < 5; j++) dx[kDiffK + j] = U(1e-5, 1e-3);
    for (int j = 0; j < 5; j++) dx[kDiffO + j] = U(1e-3, 0.5);
    dx[kDiffNut] = U(1e-4, 1e-2);
    dx[kDiffYw] = U(0.5, 40.);
    dx[kDiffRho] = U(0.99, 1.01);
    for (int j = 0; j < 4; j++) dx[kDiffDx + j] = trial % 3 == 0 ? U(0.15, 1.) : 1.;
    double e1 = check<kDiffInputs, kDiffOutputs>("koDiff", dx, [&](const auto& in, auto& out) {
      koDiffT(in, nu, out);
    });
    // rows
    double rx[kRowInputs];
    for (int j = 0; j < 10; j++) rx[j] = nu + U(0., 5e-3);
    rx[kRowMx] = U(-1e-2, 1e-2);
    rx[kRowMy] = U(-1e-2, 1e-2);
    rx[kRowUx] = U(0.0, 0.1);
    rx[kRowUy] = U(-0.01, 0.01);
    rx[kRowRho] = U(0.99, 1.01);
    rx[kRowK] = U(1e-5, 1e-3);
    rx[kRowO] = U(1e-3, 0.5);
    rx[kRowF1] = U(0.05, 0.95);
    rx[kRowNut] = U(1e-4, 1e-2);
    for (int j = 0; j < 3; j++) rx[kRowS + j] = U(-1e-3, 1e-3);
    rx[kRowYw] = U(0.5, 40.);
    for (int j = 0; j < 4; j++) rx[kRowDx + j] = trial % 3 == 0 ? U(0.15, 1.) : 1.;
    const bool nesw = trial % 4 == 1;
    double e2 = check<kRowInputs, kRowOutputs>("koRow", rx, [&](const auto& in, auto& out) {
      koRowT(in, nesw, nu, dt, out);
    });
    // eddy viscosity (one output)
    double nx[kNutInputs] = {U(1e-5, 1e-3), U(1e-3, 0.5), U(0.5, 40.), U(-1e-2, 1e-2),
                             U(-1e-2, 1e-2), U(-1e-2, 1e-2)};
    double e3 = check<kNutInputs, 1>("nut", nx, [&](const auto& in, auto& out) {
      out[0] = nutT(in, nu, 1e3);
    });
    // strain
    double gx[9];
    for (int j = 0; j < 9; j++) gx[j] = U(-1e-3, 1e-3);
    double e4 = check<9, 3>("strain", gx, [&](const auto& in, auto& out) { strainT(in, 1.25, out); });
    for (double e : {e1, e2, e3, e4})
      if (e > 1e-6) bad++;
  }
  // the difference-weight fallback and clamps are exercised by the random
  // spacings; sqrt at zero strain keeps a finite (zero) derivative
  {
    double nx[kNutInputs] = {1e-4, 0.1, 5., 0., 0., 0.};
    double yb[1] = {1.};
    double xb[kNutInputs] = {};
    transposeApply<kNutInputs, 1>(nx, yb, xb,
                                  [&](const Dual (&in)[kNutInputs], Dual (&out)[1]) {
                                    out[0] = nutT(in, 0.01, 1e3);
                                  },
                                  [](int) { return false; });
    for (double v : xb)
      if (!std::isfinite(v)) bad++;
  }
  // the adjoint kernels evaluate the non-spacing inputs with the difference
  // weights held as constants, the spacings through the full map: the same
  // transpose
  for (int trial = 0; trial < 50; trial++) {
    double rx[kRowInputs], yb[kRowOutputs], full[kRowInputs] = {}, split[kRowInputs] = {};
    for (int j = 0; j < kRowInputs; j++) rx[j] = U(0.01, 0.02);
    rx[kRowUx] = U(0., 0.1);
    rx[kRowK] = U(1e-5, 1e-3);
    rx[kRowO] = U(1e-3, 0.5);
    rx[kRowYw] = U(0.5, 40.);
    for (int j = 0; j < 4; j++) rx[kRowDx + j] = U(0.15, 1.);
    for (double& v : yb) v = U(-1., 1.);
    const bool nesw = trial % 3 == 0;
    transposeApply<kRowInputs, kRowOutputs>(
        rx, yb, full,
        [&](const Dual (&v)[kRowInputs], Dual (&o)[kRowOutputs]) { koRowT(v, nesw, 0.01, 1., o); },
        [](int) { return false; });
    double ax[6], ay[6];
    fdCoeffsT<double>(rx[kRowDx], rx[kRowDx + 1], ax);
    fdCoeffsT<double>(rx[kRowDx + 2], rx[kRowDx + 3], ay);
    transposeRange<kRowInputs, kRowOutputs>(
        rx, yb, 0, kRowDx,
        [&](const Dual (&v)[kRowInputs], Dual (&o)[kRowOutputs]) {
          koRowCore<Dual, double>(v, ax, ay, nesw, 0.01, 1., o);
        },
        [&](int i, double g) { split[i] = g; });
    transposeRange<kRowInputs, kRowOutputs>(
        rx, yb, kRowDx, kRowInputs,
        [&](const Dual (&v)[kRowInputs], Dual (&o)[kRowOutputs]) { koRowT(v, nesw, 0.01, 1., o); },
        [&](int i, double g) { split[i] = g; });
    for (int i = 0; i < kRowInputs; i++)
      if (std::fabs(full[i] - split[i]) > 1e-13 * (std::fabs(full[i]) + 1e-300)) {
        std::printf("split row transpose input %d: %.17e vs %.17e\n", i, split[i], full[i]);
        bad++;
      }
  }
  // the hand-written reverse of the row and diffusivity maps (constant
  // difference weights) against their dual-number transposes
  for (int trial = 0; trial < 400; trial++) {
    const double nu = U(0.002, 0.02);
    double ax[6], ay[6];
    fdCoeffsT<double>(trial % 3 ? 1. : U(0.15, 1.), trial % 5 ? 1. : U(0.15, 1.), ax);
    fdCoeffsT<double>(trial % 2 ? 1. : U(0.15, 1.), U(0.15, 1.), ay);
    double rx[kRowInputs], yb[kRowOutputs], dual[kRowInputs] = {}, rev[kRowInputs] = {};
    for (int j = 0; j < 10; j++) rx[j] = nu + U(0., 5e-3);
    rx[kRowMx] = U(-1e-2, 1e-2);
    rx[kRowMy] = U(-1e-2, 1e-2);
    rx[kRowUx] = U(0., 0.1);
    rx[kRowUy] = U(-0.01, 0.01);
    rx[kRowRho] = U(0.99, 1.01);
    rx[kRowK] = trial % 7 == 0 ? -1e-6 : U(1e-6, 1e-3);
    rx[kRowO] = trial % 11 == 0 ? -1e-3 : U(1e-3, 0.5);
    rx[kRowF1] = U(0.05, 0.95);
    rx[kRowNut] = U(1e-4, 1e-2);
    // strains from small (production limited) to large (unlimited), signs mixed
    const double sm = trial % 4 == 0 ? 1e-1 : 1e-3;
    for (int j = 0; j < 3; j++) rx[kRowS + j] = U(-sm, sm);
    rx[kRowYw] = trial % 6 == 0 ? U(0.01, 0.1) : U(0.5, 40.);
    for (int j = 0; j < 4; j++) rx[kRowDx + j] = 1.;
    for (double& v : yb) v = U(-1., 1.);
    const bool nesw = trial % 4 == 1 || trial % 6 == 0;
    transposeRange<kRowInputs, kRowOutputs>(
        rx, yb, 0, kRowDx,
        [&](const Dual (&v)[kRowInputs], Dual (&o)[kRowOutputs]) {
          koRowCore<Dual, double>(v, ax, ay, nesw, nu, 1., o);
        },
        [&](int i, double g) { dual[i] = g; });
    koRowCoreAdj(rx, ax, ay, nesw, nu, 1., yb, rev);
    double scale = 0.;
    for (int i = 0; i < kRowDx; i++) scale = std::fmax(scale, std::fabs(dual[i] * rx[i]));
    for (int i = 0; i < kRowDx; i++)
      if (std::fabs(dual[i] - rev[i]) * std::fabs(rx[i]) > 1e-12 * scale + 1e-300) {
        std::printf("row reverse input %d (nesw %d): %.17e vs %.17e\n", i, int(nesw), rev[i], dual[i]);
        bad++;
      }
    double dxi[kDiffInputs], dyb[kDiffOutputs], ddual[kDiffInputs] = {}, drev[kDiffInputs] = {};
    for (int j = 0; j < 5; j++) dxi[kDiffK + j] = U(1e-7, 1e-3);
    for (int j = 0; j < 5; j++) dxi[kDiffO + j] = U(1e-3, 0.5);
    dxi[kDiffNut] = U(1e-4, 1e-2);
    dxi[kDiffYw] = trial % 3 == 0 ? U(0.02, 0.5) : U(0.5, 40.);
    dxi[kDiffRho] = U(0.99, 1.01);
    for (int j = 0; j < 4; j++) dxi[kDiffDx + j] = 1.;
    for (double& v : dyb) v = U(-1., 1.);
    transposeRange<kDiffInputs, kDiffOutputs>(
        dxi, dyb, 0, kDiffDx,
        [&](const Dual (&v)[kDiffInputs], Dual (&o)[kDiffOutputs]) {
          koDiffCore<Dual, double>(v, ax, ay, nu, o);
        },
        [&](int i, double g) { ddual[i] = g; });
    koDiffCoreAdj(dxi, ax, ay, nu, dyb, drev);
    scale = 0.;
    for (int i = 0; i < kDiffDx; i++) scale = std::fmax(scale, std::fabs(ddual[i] * dxi[i]));
    for (int i = 0; i < kDiffDx; i++)
      if (std::fabs(ddual[i] - drev[i]) * std::fabs(dxi[i]) > 1e-12 * scale + 1e-300) {
        std::printf("diffusivity reverse input %d: %.17e vs %.17e\n", i, drev[i], ddual[i]);
        bad++;
      }
  }
  // the hand-written reverse of the eddy viscosity and the strain
  for (int trial = 0; trial < 400; trial++) {
    const double nu = U(0.002, 0.02);
    const double sm = trial % 3 == 0 ? 1e-5 : 1e-2;
    double nx[kNutInputs] = {trial % 9 == 0 ? 0. : U(1e-7, 1e-3), U(1e-3, 0.5),
                             trial % 4 == 0 ? U(0.01, 0.3) : U(0.5, 40.), U(-sm, sm), U(-sm, sm),
                             U(-sm, sm)};
    const double maxNut = trial % 5 == 0 ? 1e-3 : 1e3;
    const double nb = U(-1., 1.);
    double dual[kNutInputs] = {}, rev[kNutInputs] = {};
    const double ybn[1] = {nb};
    transposeRange<kNutInputs, 1>(
        nx, ybn, 0, kNutInputs,
        [&](const Dual (&v)[kNutInputs], Dual (&o)[1]) { o[0] = nutT(v, nu, maxNut); },
        [&](int i, double g) { dual[i] = g; });
    nutAdj(nx, nu, maxNut, nb, rev);
    double scale = 1e-300;
    for (int i = 0; i < kNutInputs; i++) scale = std::fmax(scale, std::fabs(dual[i] * nx[i]));
    for (int i = 0; i < kNutInputs; i++)
      if (std::fabs(dual[i] - rev[i]) * std::fmax(std::fabs(nx[i]), 1e-300) > 1e-12 * scale) {
        std::printf("nut reverse input %d: %.17e vs %.17e\n", i, rev[i], dual[i]);
        bad++;
      }
    double gx[9], sb[3], sd[9] = {}, sr[9];
    for (int j = 0; j < 9; j++) gx[j] = U(-3e-2, 3e-2);
    for (double& v : sb) v = U(-1., 1.);
    transposeRange<9, 3>(
        gx, sb, 0, 9, [&](const Dual (&v)[9], Dual (&o)[3]) { strainT(v, 1.6, o); },
        [&](int i, double g) { sd[i] = g; });
    strainAdj(gx, 1.6, sb, sr);
    double ss = 0.;
    for (int i = 0; i < 9; i++) ss = std::fmax(ss, std::fabs(sd[i]));
    for (int i = 0; i < 9; i++)
      if (std::fabs(sd[i] - sr[i]) > 1e-12 * ss) {
        std::printf("strain reverse input %d: %.17e vs %.17e\n", i, sr[i], sd[i]);
        bad++;
      }
  }
  std::printf("failures %d\n", bad);
  return bad ? 1 : 0;
}
