// engine.hpp -- device-side state and launch sequencing for one lattice.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "geometry.hpp"

namespace kofm {

struct Params {
  double tau0 = 1.25, Fval = 0., nu = 0.1, rho0 = 1., maxNut = 10., koDt = 1.;
  double relTol[2] = {1e-4, 1e-4};  // k, omega
  double absTol[2] = {1e-7, 1e-7};
  int maxIters[2] = {1000, 1000};
};

struct SolverStats {
  int iters[2] = {0, 0};
  double res0[2] = {0., 0.}, res[2] = {0., 0.};
  int nan = 0;          // non-finite residuals since the last query
  int stepsPerLB = 1;   // k-omega sub-steps per LB step
  double timeStep = 1.; // their time step
  int batches = 0;      // batches of k-omega steps run from one graph since the last query
  int rollbacks = 0;    // of those, batches rerun one step at a time
  int tIters = 0;       // temperature: iterations of the last solve, its residuals,
  double tRes0 = 0., tRes = 0.;  // sub-steps per call and their time step
  int tStepsPerLB = 0;
  double tTimeStep = 0.;
};

// The temperature solve (clVariablesFD): diffusivities, turbulent Prandtl
// number, inlet and initial temperatures, the time step dTtfd and the
// solver tolerances with BiCGStabSolver's meaning.
struct ThermalParams {
  double alphaFluid = 0., alphaFoul = 0., prTurb = 0.9, tIn = 2., tInitial = 0.5, dt = 1.;
  double relTol = 1e-4, absTol = 1e-7;
  int maxIters = 1000;
};

// The Nusselt-number output (clVariablesFD::iniNuCoeffs): the rows the bulk
// temperature sums, the wavy section whose period boundaries get column sums,
// and the conductivities of soot and air in lattice units.  The columns are
// the particle domain of the built geometry, floor(X Min Val) to
// floor(X Stop Pos).
struct NusseltParams {
  int height = 0, wavyStart = -1, wavyPeriod = -1, wavyPeriods = -1;
  double kSoot = 0., kAir = 0.;
};

// TR_shear_removal's constants (particleProperties' arrays and the
// tracer's defines) and the tracer's particle arrays (device pointers).
struct RemovalParams {
  int nSizes = 0;
  const double *tauCrit = nullptr, *dcoeff = nullptr, *lcoeff = nullptr, *mp = nullptr;
  double dtWall = 3., foulSwitch = 0.5;
  int locStride = 0, timerStart = 0;
};
struct ParticleArrays {
  double* pos = nullptr;
  const uint32_t* numRep = nullptr;
  const int16_t* type = nullptr;
  int32_t* depFlag = nullptr;
  uint16_t* depTimer = nullptr;
  int32_t* loc = nullptr;
  int32_t* timer = nullptr;
};

// J = mean over the last `window` steps of a run of
//   pbar(xIn) - pbar(xOut) + shearWeight * sum over shear links of tau_BL * |link|
// with pbar the mean of rho / 3 over the fluid nodes of a column.
struct Objective {
  int window = 1, xIn = 0, xOut = 0;
  int flags = 0;  // KOFM_ADJOINT_*
  double shearWeight = 0.;
};

// Type-erased engine; the implementation is templated on the population
// storage type and the arithmetic type.
class Engine {
 public:
  virtual ~Engine() = default;
  virtual int uploadGeometry(const Geometry& g, std::string& err) = 0;
  // Moves the walls to the points C (the built geometry's points and links)
  // and updates the node types, links, spacings, wall distances and
  // wall-shear tables; the updated geometry is returned by geometry().
  virtual int updateWalls(const std::vector<double>& C, std::string& err) = 0;
  virtual const Geometry& geometry() const = 0;
  virtual void setParams(const Params& p) = 0;
  virtual int initState(const double* rho, const double* ux, const double* uy,
                        std::string& err) = 0;
  virtual int initKOmega(double k, double w, double Re, double geomPenalty, double pipeRadius,
                         std::string& err) = 0;
  virtual int uploadState(const double* f, const double* k, const double* w, const double* nut,
                          const double* S, std::string& err) = 0;
  virtual int step(int n, unsigned options, std::string& err) = 0;
  // msKernel null: the steps run as step() does, timed as a whole; otherwise
  // events also bracket every collide-stream launch (their own cost then
  // lands in msStep too).
  virtual int timeSteps(int n, unsigned options, double& msStep, double* msKernel,
                        std::string& err) = 0;
  virtual int wallShear(std::string& err) = 0;
  virtual int reduce(int field, int op, double& out, std::string& err) = 0;
  virtual int download(int field, double* host, std::string& err) = 0;
  virtual int solverStats(SolverStats& s, std::string& err) = 0;
  virtual int synchronize(std::string& err) = 0;
  // KOFM_DT_* action on the k-omega time step; returns the sub-steps per LB
  // step and their time step after it.
  virtual int komegaTimeStep(int action, int& stepsPerLB, double& dt, std::string& err) = 0;
  // clVariablesFD::ini, ::Solve (n calls), halveTimeStep/doubleTimeStep, and
  // the temperature field
  virtual int thermalInit(const ThermalParams& q, std::string& err) = 0;
  virtual int thermalStep(int n, std::string& err) = 0;
  virtual int thermalTimeStep(int action, int& stepsPerLB, double& dt, std::string& err) = 0;
  virtual int setTemperature(const double* T, std::string& err) = 0;
  virtual bool thermalOn() const = 0;
  // clVariablesFD::iniNuCoeffs and ::updateTimeData: per column 5 values
  // (bulk temperature, then gradient and wall temperature of the bottom and
  // top walls), per period boundary 2 (sum |u| T, sum |u|); optionally the
  // nearest node of each (column, wall) slot, (row << 2) | type, and its
  // vector from the wall.
  virtual int nusseltInit(const NusseltParams& q, std::string& err) = 0;
  virtual int nusseltSizes(int& columns, int& locations, std::string& err) const = 0;
  virtual int nusselt(double* nu, double* mean, int32_t* nodes, double* vec, std::string& err) = 0;
  // TR_shear_removal over particles [offset, offset + count), enqueued on
  // the context stream with the last computed link shear
  virtual int shearRemoval(const RemovalParams& q, const ParticleArrays& P, int64_t offset,
                           int64_t count, uint32_t* blDep, std::string& err) = 0;
  // the tracer's link arrays the removal uses (first point, unit normal,
  // length, surface type), nBL each, for the walls now
  virtual int removalLinks(double foulSwitch, double* p0, double* vN, double* len, int16_t* surf,
                           std::string& err) = 0;
  virtual int64_t fieldSize(int field) const = 0;
  virtual int setLaunch(int block, int npt, int spp, int64_t pad, std::string& err) = 0;
  virtual void setGraphs(bool on) = 0;
  virtual int autotune(int steps, int& block, int& npt, int& spp, int64_t& pad, double& ms,
                       std::string& err) = 0;
  // Runs n steps accumulating J and keeping a population checkpoint every
  // `interval` steps; adjoint() then returns dJ/dy of the wall points (and
  // dJ/dq of the boundary links, if dJdq is not null).
  virtual int objectiveRun(int n, int interval, const Objective& o, double& J,
                           std::string& err) = 0;
  virtual int adjoint(double* dJdy, double* dJdq, std::string& err) = 0;
};

class Ensemble {
 public:
  virtual ~Ensemble() = default;
  // A member's destructor calls this; every later call on the ensemble then
  // fails with KOFM_ERR_STATE.
  virtual void memberGone(const Engine* e) = 0;
  virtual int step(int n, unsigned options, std::string& err) = 0;
  virtual int timeSteps(int n, unsigned options, double& ms, std::string& err) = 0;
  virtual int setBlock(int block, std::string& err) = 0;
  virtual int setLaunch(int block, int spp, std::string& err) = 0;
  virtual int setCacheBudget(int64_t bytes, std::string& err) = 0;
  virtual int autotune(int steps, int& block, int& spp, int64_t& budget, double& ms,
                       std::string& err) = 0;
  virtual int groups() const = 0;
  virtual int synchronize(std::string& err) = 0;
  virtual int64_t fluidNodes() const = 0;
};

// Error text returned by kofm_last_error(NULL) (calls without a context).
void setGlobalError(const std::string& msg);

// storage: KOFM_STORE_*
// rc receives the negative KOFM_ERR_* code of a failure
std::unique_ptr<Engine> makeEngine(int nX, int nY, int storage, bool komega, int& rc,
                                   std::string& err);
std::unique_ptr<Ensemble> makeEnsemble(const std::vector<Engine*>& members, std::string& err);

}  // namespace kofm
