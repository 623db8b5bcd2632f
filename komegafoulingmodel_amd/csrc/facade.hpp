// facade.hpp -- the reference's host classes for the lattice-Boltzmann step
// (clProblem, clVariablesLS, clVariablesLB with its kOmega member,
// shearStress, Reducer), re-expressed over the C ABI of include/kofm.h.  A
// host program written against the reference's call sequence
// (LBLS.cpp:65-160 ini, 185-381 run) drives the MI355X path through these
// classes; they own no device code and call nothing but kofm_*.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <climits>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/kofm.h"
#include "runparams.hpp"

namespace kofm {
namespace ref {

inline void check(kofm_ctx* c, int rc, const char* what) {
  if (rc != KOFM_OK)
    throw std::runtime_error(std::string(what) + ": " + kofm_last_error(c) + " (code " +
                             std::to_string(rc) + ")");
}

// clProblem::loadParams (clProblem.cpp:456-554): sizes and cadences.
class clProblem {
 public:
  int nX = 0, nY = 0, XsizeFull = 0, Channel_Height = 0, DeviceID = 0;
  double Pipe_radius = 0.;
  unsigned StopTime = 0, Time = 0, RestartTime = 0;
  unsigned trSteps_wall = 3;
  unsigned tfdSteps = 1;  // "LB Steps Per T" (clProblem.cpp:506); dTtfd = tfdSteps
  double dTtfd = 1.;
  unsigned saveAvgStartTime = 0, saveTimeStepAvg = 1000, dumpStartTime = 0, dumpTimeStep = 1000;
  unsigned saveNuStartTime = 0, saveTimeStepNu = 1000;
  // the wavy section (clProblem.cpp:471-473), -1 when not given
  int xWavyStart = -1, wavyPeriodLen = -1, numWavyPeriods = -1;
  unsigned saveBinStep = 1;  // dumps per restart-file save ("Save Bin Step", clProblem.cpp:515)
  // The cadence state a restart resumes from (clProblem.cpp:540-550, written
  // by saveParameters :385-403): the next time-data row and dump, the dumps
  // so far and the dump count at which the restart files are next written.
  unsigned nextSaveAvgTime = 0, nextSaveNuTime = 0, nextDumpStepTime = 0;
  int currentDumpStep = 0, nextDumpBin = 1;
  RunParams params;

  void loadRunParameters(const std::string& path) {
    if (!params.load(path)) throw std::runtime_error("cannot read " + path);
    const int d = RunParams::kSystem;
    nX = params.getInt(d, "nX", 0);
    nY = params.getInt(d, "nY", 0);
    if (nX <= 0 || nY <= 0) throw std::runtime_error("nX and nY are required");
    XsizeFull = int(std::ceil(double(nX) / 128.)) * 128;
    Pipe_radius = params.getDouble(d, "Pipe Radius", 0.5 * (nY - 4));
    Channel_Height = 2 * int(Pipe_radius);
    DeviceID = params.getInt(d, "Device ID", 0);
    RestartTime = unsigned(params.getInt(d, "Current Time", 0));
    Time = RestartTime;
    StopTime = unsigned(params.getInt(d, "Stop Time", 200000000));
    trSteps_wall = unsigned(params.getInt(d, "LB Steps Per TR Wall", 3));
    const int tfd = params.getInt(d, "LB Steps Per T", 1);
    if (tfd < 1) throw std::runtime_error("LB Steps Per T must be >= 1");
    tfdSteps = unsigned(tfd);
    dTtfd = double(tfdSteps);
    xWavyStart = params.getInt(d, "Wavy Start Loc", -1);
    wavyPeriodLen = params.getInt(d, "Period Length", -1);
    numWavyPeriods = params.getInt(d, "Num Periods", -1);
    cadence("Save Avgs", 2000, saveAvgStartTime, saveTimeStepAvg);
    cadence("Save Nu", 2000, saveNuStartTime, saveTimeStepNu);
    cadence("Dump", 1000, dumpStartTime, dumpTimeStep);
    const int bin = params.getInt(d, "Save Bin Step", 1);
    saveBinStep = unsigned(bin > 0 ? bin : 1);
    nextSaveAvgTime =
        unsigned(params.getInt(d, "nextSaveAvgTime", int(std::max(saveAvgStartTime, saveTimeStepAvg))));
    nextSaveNuTime =
        unsigned(params.getInt(d, "nextSaveNuTime", int(std::max(saveNuStartTime, saveTimeStepNu))));
    nextDumpStepTime = unsigned(params.getInt(d, "nextDumpStepTime", int(dumpStartTime + dumpTimeStep)));
    currentDumpStep = params.getInt(d, "currentDumpStep", 0);
    nextDumpBin = params.getInt(d, "nextDumpBin", int(saveBinStep));
  }

  // clProblem::saveParameters (clProblem.cpp:256-405): RunParam.yaml with the
  // current time, marked as a restart, into dir.
  void saveParameters(const std::string& dir) {
    const int d = RunParams::kSystem;
    params.setInt(d, "Current Time", int(Time));
    params.setBool(d, "Restart Run", true);
    params.setInt(d, "nextSaveAvgTime", int(nextSaveAvgTime));
    params.setInt(d, "nextSaveNuTime", int(nextSaveNuTime));
    params.setInt(d, "nextDumpStepTime", int(nextDumpStepTime));
    params.setInt(d, "currentDumpStep", currentDumpStep);
    params.setInt(d, "nextDumpBin", nextDumpBin);
    if (!params.save(dir + "/RunParam.yaml")) throw std::runtime_error("cannot write " + dir + "/RunParam.yaml");
  }

 private:
  // getSaveStepInfo (clProblem.cpp:430-454)
  void cadence(const std::string& name, unsigned defNum, unsigned& start, unsigned& step) {
    const int d = RunParams::kSystem;
    start = unsigned(params.getInt(d, name + " Start Time", 0));
    double num = 0., ts = 0.;
    const int which = params.either(d, name + " Num Steps", name + " Timestep", num, ts);
    if (which == 2)
      step = unsigned(ts);
    else
      step = (StopTime - start) / unsigned(which == 1 ? num : defNum);
    if (step == 0) step = 1;
  }
};

// Reducer<...>::reduceSingle / reduce / appendToFileFromDevice
// (Reducer.h:88-290) for one field and operator, with the TimeData text
// file "<name>.txt" (time, value rows; Array.h:2818-2904): reduce() keeps
// one row per call and appends the rows to the file by itself whenever
// maxLines of them are held (setTimeAndIncrement; the reference holds
// OUTPUT_MAX_LINES = 128, ConstDef.h:320); appendToFileFromDevice() flushes
// the rows held so far.
class Reducer {
 public:
  Reducer() = default;
  void ini(kofm_ctx* ctx, int field, int op, const std::string& name, const std::string& dir = ".",
           int maxLines = 128) {
    ctx_ = ctx;
    field_ = field;
    op_ = op;
    name_ = name;
    dir_ = dir;
    maxLines_ = maxLines > 0 ? maxLines : 1;
  }
  void setDir(const std::string& dir) { dir_ = dir; }
  double reduceSingle() {
    double v = 0.;
    check(ctx_, kofm_reduce(ctx_, field_, op_, &v), "reduce");
    return v;
  }
  // returns true when the rows were written out (the buffer was full)
  bool reduce(unsigned time) {
    rows_.push_back({double(time), reduceSingle()});
    if (int(rows_.size()) < maxLines_) return false;
    appendToFileFromDevice();
    return true;
  }
  void appendToFileFromDevice() {
    if (rows_.empty()) return;
    std::ofstream out(dir_ + "/" + name_ + ".txt", std::ios::app);
    for (auto& r : rows_) out << r[0] << "\t" << r[1] << "\n";
    rows_.clear();
  }
  size_t held() const { return rows_.size(); }

 private:
  kofm_ctx* ctx_ = nullptr;
  int field_ = 0, op_ = KOFM_RED_SUM;
  int maxLines_ = 128;
  std::string name_, dir_ = ".";
  std::vector<std::vector<double>> rows_;
};

// TimeData<double> (Array.h:2702-2904) for rows of several values: the
// file "<name>.txt" is emptied at the start of a run that is not a restart
// (iniFile), rows are held and written out "time\tv0\tv1\t...\t" when
// maxLines are held (setTimeAndIncrement) or on appendData.
class TimeRows {
 public:
  void ini(const std::string& name, const std::string& dir, int maxLines, bool restart) {
    path_ = dir + "/" + name + ".txt";
    maxLines_ = maxLines > 0 ? maxLines : 1;
    rows_.clear();
    if (!restart) {
      std::ofstream out(path_, std::ios::trunc);
      if (!out) throw std::runtime_error("cannot create " + path_);
    }
  }
  void add(unsigned time, std::vector<double> row) {
    times_.push_back(time);
    rows_.push_back(std::move(row));
    if (int(rows_.size()) >= maxLines_) appendData();
  }
  void appendData() {
    if (rows_.empty()) return;
    std::ofstream out(path_, std::ios::app);
    for (size_t r = 0; r < rows_.size(); r++) {
      out << times_[r] << "\t";
      for (double v : rows_[r]) out << v << "\t";
      out << "\n";
    }
    rows_.clear();
    times_.clear();
  }

 private:
  std::string path_;
  int maxLines_ = 10;
  std::vector<unsigned> times_;
  std::vector<std::vector<double>> rows_;
};

// clVariablesLS (clVariablesLS.cpp:487-620): wall polyline in, node types
// and boundary lists built (here by kofm_build_geometry).
class clVariablesLS {
 public:
  int nN = 0, nBL = 0;
  std::vector<double> C0;
  std::vector<int32_t> BL;
  double xReleasePos = 90., xMinVal = 88.75, xStopPos = 1900., cutoffRadius = 5.;
  int indRadiusSearch = 30, wallDSearchRad = 50;
  double minDxDist = 0.001;

  void loadParams(const clProblem& p, const std::string& loadDir = "load") {
    const RunParams& rp = p.params;
    nN = rp.getInt(RunParams::kLevelSet, "nN", 0);
    nBL = nN - 2;
    // geometry inputs that the reference reads in other classes
    xReleasePos = rp.getDouble(RunParams::kParticles, "X Release Pos", 90.);
    xMinVal = rp.getDouble(RunParams::kParticles, "X Min Val", xReleasePos - 1.25);
    xStopPos = rp.getDouble(RunParams::kParticles, "X Stop Pos", 1900.);
    cutoffRadius = rp.getDouble(RunParams::kParticles, "Cutoff Radius", 5.);
    indRadiusSearch = rp.getInt(RunParams::kParticles, "Index Radius Search", 30);
    wallDSearchRad = rp.getInt(RunParams::kFluid, "Wall D Search Rad", 50);
    minDxDist = rp.getDouble(RunParams::kFouling, "Min dX Dist", 0.001);
    // load/lsc0.txt: "x y" rows; load/lsbl.txt: index pairs (allocateArrays, :24-51)
    C0 = readTable<double>(loadDir + "/lsc0.txt", 2);
    BL = readTable<int32_t>(loadDir + "/lsbl.txt", 2);
    if (nN == 0) nN = int(C0.size() / 2);
    if (int(C0.size() / 2) != nN) throw std::runtime_error("lsc0.txt does not hold nN points");
    nBL = int(BL.size() / 2);
  }

  void ini(kofm_ctx* ctx, const clProblem& p) {
    kofm_geom_params gp{};
    gp.pipe_radius = p.Pipe_radius;
    gp.eps = 1e-5;
    gp.min_dx = minDxDist;
    gp.wall_d_search_rad = wallDSearchRad;
    gp.x_release_pos = xReleasePos;
    gp.x_min_val = xMinVal;
    gp.x_stop_pos = xStopPos;
    gp.cutoff_radius = cutoffRadius;
    gp.index_radius_search = indRadiusSearch;
    check(ctx, kofm_build_geometry(ctx, C0.data(), nN, BL.data(), nBL, &gp), "geometry");
  }

  template <class T>
  static std::vector<T> readTable(const std::string& path, int cols) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot read " + path);
    std::vector<T> v;
    double x;
    while (in >> x) v.push_back(T(x));
    if (v.size() % cols) throw std::runtime_error(path + " has a ragged last row");
    return v;
  }
};

// kOmega (kOmega.h:223-330): parameters and initial values; its transport
// update runs inside kofm_step.
class kOmega {
 public:
  bool kOmegaSolverFlag = false;
  double kappaMaxRelTol = 1e-4, kappaMaxAbsTol = 1e-7, omegaMaxRelTol = 1e-4, omegaMaxAbsTol = 1e-7;
  int kappaMaxIters = 1000, omegaMaxIters = 1000;
  double maxTurbVisc = 10., kIniVal = -1., omegaIniVal = -1.;
  // initial velocity field of iniDists(Umax) (ConstDef.h:159-190 defaults)
  bool iniTurbVelocity = true, perturbVelField = true;
  double perturbDUPlus = 0.025, perturbEpsilon = 250., perturbBeta = 100., perturbAlpha = 55.75,
         perturbSigma = 0.00055;
  double roughnessFactor = 9.;  // read only: its use is commented out (kOmega.cpp:736-754)
  kofm_ctx* ctx = nullptr;

  // kOmega::loadParams (kOmega.cpp:352-378).  The reference reads the omega
  // initial value from the key "K Initial Val" (kOmega.cpp:376); here it is
  // read from "Omega Initial Val", the key kOmega::saveParams writes.
  void loadParams(const RunParams& rp) {
    const int d = RunParams::kFluid;
    kOmegaSolverFlag = rp.getBool(d, "Use Turb Model", true);
    kappaMaxRelTol = rp.getDouble(d, "Kappa Max Rel Tol", 1e-4);
    kappaMaxAbsTol = rp.getDouble(d, "Kappa Max Abs Tol", 1e-7);
    kappaMaxIters = rp.getInt(d, "Kappa Max Iterations", 1000);
    omegaMaxRelTol = rp.getDouble(d, "Omega Max Rel Tol", kappaMaxRelTol);
    omegaMaxAbsTol = rp.getDouble(d, "Omega Max Abs Tol", kappaMaxAbsTol);
    omegaMaxIters = rp.getInt(d, "Omega Max Iterations", kappaMaxIters);
    maxTurbVisc = rp.getDouble(d, "Max Turb Visc", 10.);
    kIniVal = rp.getDouble(d, "K Initial Val", -1.);
    omegaIniVal = rp.getDouble(d, "Omega Initial Val", -1.);
    iniTurbVelocity = rp.getBool(d, "Ini Turb Velocity", true);
    perturbVelField = rp.getBool(d, "Perturb Vel Field", true);
    perturbDUPlus = rp.getDouble(d, "DUplus Perturb Multiplier", 0.025);
    perturbEpsilon = rp.getDouble(d, "Epsilon Perturb Multiplier", 250.);
    perturbBeta = rp.getDouble(d, "Beta Perturb Multiplier", 100.);
    perturbAlpha = rp.getDouble(d, "Alpha Perturb Multiplier", 55.75);
    perturbSigma = rp.getDouble(d, "Sigma Perturb Multiplier", 0.00055);
    roughnessFactor = rp.getDouble(d, "Wall Roughness Factor", 9.);
  }
  void ini(kofm_ctx* c, double Re, double geomPenaltyFactor) {
    ctx = c;
    if (kOmegaSolverFlag)
      check(ctx, kofm_init_komega(ctx, kIniVal, omegaIniVal, Re, geomPenaltyFactor), "k-omega init");
  }
  void Solve() {}  // fused into clVariablesLB::Solve
  // kOmega::halveTimeStep / doubleTimeStep (kOmega.cpp:387-414); kofm_step
  // calls them itself on a non-finite residual and after 20 quiet steps
  void halveTimeStep() { check(ctx, kofm_komega_time_step(ctx, KOFM_DT_HALVE, nullptr, nullptr), "halve"); }
  void doubleTimeStep() { check(ctx, kofm_komega_time_step(ctx, KOFM_DT_DOUBLE, nullptr, nullptr), "double"); }
};

// clVariablesLB (clVariablesLB.h:40-458): fluid parameters, initial state,
// the time step, output.
class clVariablesLB {
 public:
  double tau = 1.99, MuVal = 0., Re = 10800., Fval = 0., RhoVal = 1.;
  double nuAir = 4.1e-5;  // "Nu air": DELTA_L^2 / DELTA_T = MuVal / nuAir (clVariablesLB.cpp:608)
  double rhoAir = 0.675;  // "Rho air" (DENSITY_AIR, ConstDef.h:60): DELTA_M (clVariablesLB.cpp:609)
  bool tauDefined = true, reDefined = true;
  int tlbmIniDumpStep = 100000, tlbmNumIniSteps_LB = 1000000;
  // ConstDef.h:159-192 defaults
  bool saveAvgInfo = true, loadVelTxtFlag = true, iniPoiseuille = true, saveMacroStart = false;
  bool turbVelBCFlag = true;  // accepted; its kernel code is commented out (lbKernels.cl:428-460)
  double geomPenaltyFactor = 1.;
  int lbOutSize = 128;        // TimeData rows kept before a flush ("lbOut Max Lines")
  int storage = KOFM_STORE_F32;
  kOmega kOmegaClass;
  Reducer sumUx, sumUy, sumRo, sumKappa, sumOmega, minKappa, minOmega;
  kofm_ctx* ctx = nullptr;

  ~clVariablesLB() {
    if (ctx) kofm_destroy(ctx);
  }

  // Fluid-document keys the reference reads or writes (clVariablesLB.cpp:553-639,
  // 652-700; kOmega.cpp:352-378, 480-520).  Any other key there is an error:
  // it would otherwise be ignored without notice.
  static bool knownFluidKey(const std::string& k) {
    static const char* const keys[] = {
        "ParamsName", "Use Turb Model", "Save Avg Info", "lbOut Max Lines", "load Vel Text", "Mu",
        "Tau", "Re", "Fval", "Rho", "Ux inlet", "Nu air", "Rho air", "Save Macros On Start",
        "Ini Poiseuille", "Run LBFD First", "Ini Dump Step", "Ini LB Steps", "Ini TLBM Steps",
        "Num Intervals Per Avg", "Time Between Intervals", "Pause Between Adjust",
        "Flowrate Max Percent Diff", "Use Turb Vel BC", "Geometric Penalty Factor",
        "Ini Turb Velocity", "Perturb Vel Field", "DUplus Perturb Multiplier",
        "Epsilon Perturb Multiplier", "Beta Perturb Multiplier", "Alpha Perturb Multiplier",
        "Sigma Perturb Multiplier", "Wall Roughness Factor", "Kappa Max Rel Tol",
        "Kappa Max Abs Tol", "Kappa Max Iterations", "Omega Max Rel Tol", "Omega Max Abs Tol",
        "Omega Max Iterations", "Wall D Search Rad", "K Initial Val", "Omega Initial Val",
        "Max Turb Visc", "Restart Run",
        "Population Storage"};  // no reference counterpart: KOFM_STORE_* by name
    for (const char* s : keys)
      if (k == s) return true;
    return false;
  }

  // clVariablesLB::loadParams (clVariablesLB.cpp:553-639)
  void loadParams(const clProblem& p) {
    const RunParams& rp = p.params;
    const int d = RunParams::kFluid;
    for (const std::string& k : rp.keys(d))
      if (!knownFluidKey(k))
        throw std::runtime_error("unsupported fluid parameter \"" + k + "\" in RunParam.yaml");
    // the combined LB / thermal warm-up of clVariablesLB::ini is not built
    if (rp.getBool(d, "Run LBFD First", false))
      throw std::runtime_error("\"Run LBFD First\" (the combined LB and thermal warm-up) is "
                               "not part of this build");
    nuAir = rp.getDouble(d, "Nu air", 4.1e-5);  // VISC_AIR, ConstDef.h:59
    rhoAir = rp.getDouble(d, "Rho air", 0.675);
    double mu = 0., t = 0.;
    switch (rp.either(d, "Mu", "Tau", mu, t)) {
      case 1: MuVal = mu; tau = 1. / (MuVal * 3. + 0.5); tauDefined = false; break;
      case 2: tau = t; MuVal = (1. / tau - 0.5) / 3.; break;
      default: tau = 1.99; MuVal = (1. / tau - 0.5) / 3.; break;
    }
    double re = 0., fv = 0.;
    const double R = p.Pipe_radius;
    switch (rp.either(d, "Re", "Fval", re, fv)) {
      case 1: Re = re; Fval = Re * 3. * MuVal * MuVal / R / R / R; break;
      case 2: Fval = fv; Re = Fval * R * R * R / 3. / MuVal / MuVal; reDefined = false; break;
      default: Re = 10800.; Fval = Re * 3. * MuVal * MuVal / R / R / R; break;
    }
    RhoVal = rp.getDouble(d, "Rho", 1.);
    tlbmIniDumpStep = rp.getInt(d, "Ini Dump Step", 100000);
    tlbmNumIniSteps_LB = rp.getInt(d, "Ini LB Steps", 1000000);
    saveAvgInfo = rp.getBool(d, "Save Avg Info", true);
    loadVelTxtFlag = rp.getBool(d, "load Vel Text", true);
    iniPoiseuille = rp.getBool(d, "Ini Poiseuille", true);
    saveMacroStart = rp.getBool(d, "Save Macros On Start", false);
    turbVelBCFlag = rp.getBool(d, "Use Turb Vel BC", true);
    geomPenaltyFactor = rp.getDouble(d, "Geometric Penalty Factor", 1.);
    lbOutSize = rp.getInt(d, "lbOut Max Lines", 128);
    if (lbOutSize < 1) throw std::runtime_error("lbOut Max Lines must be >= 1");
    const std::string st = rp.getString(d, "Population Storage", "f32");
    if (st == "f32") storage = KOFM_STORE_F32;
    else if (st == "f16") storage = KOFM_STORE_F16;
    else if (st == "bf16") storage = KOFM_STORE_BF16;
    else if (st == "f64") storage = KOFM_STORE_F64;
    else if (st == "f32x") storage = KOFM_STORE_F32X;
    else throw std::runtime_error("Population Storage must be f32, f16, bf16, f64 or f32x");
    kOmegaClass.loadParams(rp);
  }

  // allocation + geometry + initial state (clVariablesLB::ini, :307-365);
  // restart from load/lbf.bin (+ lbkappa.bin, lbomega.bin) when present
  void ini(const clProblem& p, clVariablesLS& ls, const std::string& loadDir = "load",
           const std::string& outDir = ".") {
    kofm_grid g{p.nX, p.nY, storage, kOmegaClass.kOmegaSolverFlag ? 1 : 0};
    if (kofm_create(&g, &ctx) != KOFM_OK) throw std::runtime_error(kofm_last_error(nullptr));
    kofm_params q{};
    q.tau0 = tau;
    q.Fval = Fval;
    q.nu = MuVal;
    q.rho0 = RhoVal;
    q.max_turb_visc = kOmegaClass.maxTurbVisc;
    q.ko_dt = 1.;
    q.kappa_rel_tol = kOmegaClass.kappaMaxRelTol;
    q.kappa_abs_tol = kOmegaClass.kappaMaxAbsTol;
    q.kappa_max_iters = kOmegaClass.kappaMaxIters;
    q.omega_rel_tol = kOmegaClass.omegaMaxRelTol;
    q.omega_abs_tol = kOmegaClass.omegaMaxAbsTol;
    q.omega_max_iters = kOmegaClass.omegaMaxIters;
    check(ctx, kofm_set_params(ctx, &q), "params");
    ls.ini(ctx, p);
    const size_t D = size_t(p.XsizeFull) * p.nY;
    std::vector<double> f;
    restarted = readBin(loadDir + "/lbf.bin", 9 * D, f);
    if (restarted) {
      std::vector<double> k, w;
      const bool ko = kOmegaClass.kOmegaSolverFlag && readBin(loadDir + "/lbkappa.bin", D, k) &&
                      readBin(loadDir + "/lbomega.bin", D, w);
      check(ctx, kofm_upload_state(ctx, f.data(), ko ? k.data() : nullptr, ko ? w.data() : nullptr,
                                   nullptr, nullptr), "restart");
      kOmegaClass.ctx = ctx;
    } else {
      // clVariablesLB::ini (clVariablesLB.cpp:311-333): velocity text files,
      // else the Poiseuille / law-of-the-wall field, else rest
      std::vector<double> ro(D, 0.), ux(D, 0.), uy(D, 0.);
      const std::string rdir = restartDir(loadDir);
      if (loadVelTxtFlag && readTxt2D(rdir + "/lbux.txt", p, ux) &&
          readTxt2D(rdir + "/lbuy.txt", p, uy) && readTxt2D(rdir + "/lbro.txt", p, ro)) {
        check(ctx, kofm_init_state(ctx, ro.data(), ux.data(), uy.data()), "iniDistsFromVel");
      } else if (iniPoiseuille) {
        kofm_ini_flow q{};
        q.Re = Re;
        q.nu = MuVal;
        q.pipe_radius = p.Pipe_radius;
        q.rho = RhoVal;
        q.turbulent_profile = kOmegaClass.iniTurbVelocity ? 1 : 0;
        q.perturb = kOmegaClass.perturbVelField ? 1 : 0;
        q.duplus_mult = kOmegaClass.perturbDUPlus;
        q.epsilon_mult = kOmegaClass.perturbEpsilon;
        q.beta_mult = kOmegaClass.perturbBeta;
        q.alpha_mult = kOmegaClass.perturbAlpha;
        q.sigma_mult = kOmegaClass.perturbSigma;
        q.seed = 1;  // clProblem.cpp:80
        check(ctx, kofm_ini_poiseuille(ctx, &q, ro.data(), ux.data(), uy.data()), "iniDists(Umax)");
        check(ctx, kofm_init_state(ctx, ro.data(), ux.data(), uy.data()), "iniDists(Umax)");
      } else {
        check(ctx, kofm_init_state(ctx, nullptr, nullptr, nullptr), "iniDists");
      }
      kOmegaClass.ini(ctx, Re, geomPenaltyFactor);
    }
    sumUx.ini(ctx, KOFM_FIELD_UX, KOFM_RED_SUM, "redUx", outDir, lbOutSize);
    sumUy.ini(ctx, KOFM_FIELD_UY, KOFM_RED_SUM, "redUy", outDir, lbOutSize);
    sumRo.ini(ctx, KOFM_FIELD_RHO, KOFM_RED_SUM, "redRo", outDir, lbOutSize);
    if (kOmegaClass.kOmegaSolverFlag) {
      sumKappa.ini(ctx, KOFM_FIELD_KAPPA, KOFM_RED_SUM, "redKappa", outDir, lbOutSize);
      sumOmega.ini(ctx, KOFM_FIELD_OMEGA, KOFM_RED_SUM, "redOmega", outDir, lbOutSize);
      minKappa.ini(ctx, KOFM_FIELD_KAPPA, KOFM_RED_MIN, "minKappa", outDir, lbOutSize);
      minOmega.ini(ctx, KOFM_FIELD_OMEGA, KOFM_RED_MIN, "minOmega", outDir, lbOutSize);
    }
  }

  // clVariablesLB::Solve (clVariablesLB.h:375-441)
  void Solve(int n = 1) { check(ctx, kofm_step(ctx, n, KOFM_OPTION_SAVE_MACRO_FIELDS), "step"); }

  double calcUmean() { return sumUx.reduceSingle() / sumRo.reduceSingle(); }

  // warm-up loop (clVariablesLB::getInitialFlow, :170-232)
  void getInitialFlow(const clProblem& p, const std::string& outDir) {
    int time = 0;
    const double sizedom = double(p.nX) * double(p.Channel_Height);
    while (time < tlbmNumIniSteps_LB) {
      const int n = std::min(1000 - time % 1000, tlbmNumIniSteps_LB - time);
      Solve(n);
      time += n;
      if (time % tlbmIniDumpStep == 0) save2file(outDir);
      if (time % 1000 == 0) {
        const double ro = sumRo.reduceSingle();
        const double ux = sumUx.reduceSingle() / ro;
        if (kOmegaClass.kOmegaSolverFlag)
          std::printf("Time = %d, Umean = %g, K = %g, Omega = %g, Re = %g\n", time, ux,
                      sumKappa.reduceSingle() / ro, sumOmega.reduceSingle() / ro,
                      ux * p.Pipe_radius / MuVal);
        else
          std::printf("Time = %d, Umean = %g, Re = %g\n", time, ux, ux * p.Pipe_radius / MuVal);
        (void)sizedom;
      }
    }
    save2file(outDir);
  }

  void save2file(const std::string& dir) { check(ctx, kofm_save2file(ctx, dir.c_str()), "save2file"); }
  // clVariablesLB/kOmega/clVariablesLS::saveRestartFiles: lbf.bin,
  // lbkappa.bin, lbomega.bin, lsc.bin
  void saveRestartFiles(const std::string& dir) {
    check(ctx, kofm_save_restart(ctx, dir.c_str()), "saveRestartFiles");
  }
  void saveDebug(const std::string& dir) { check(ctx, kofm_save_debug(ctx, dir.c_str()), "saveDebug"); }

  // clVariablesLB::updateTimeData / saveTimeData (clVariablesLB.cpp:768-790,
  // 926-945): only with "Save Avg Info"
  void updateTimeData(unsigned time) {
    if (!saveAvgInfo) return;
    sumUx.reduce(time);
    sumUy.reduce(time);
    sumRo.reduce(time);
    if (kOmegaClass.kOmegaSolverFlag) {
      sumKappa.reduce(time);
      sumOmega.reduce(time);
    }
  }
  void saveTimeData() {
    if (!saveAvgInfo) return;
    for (Reducer* r : {&sumUx, &sumUy, &sumRo}) r->appendToFileFromDevice();
    if (kOmegaClass.kOmegaSolverFlag)
      for (Reducer* r : {&sumKappa, &sumOmega}) r->appendToFileFromDevice();
  }

  bool restarted = false;  // the state came from load/lbf.bin

  // "restart_files" beside the load directory (the reference opens
  // restart_files\\lbux relative to its working directory, which holds load\\)
  static std::string restartDir(const std::string& loadDir) {
    const size_t s = loadDir.find_last_of('/');
    return (s == std::string::npos ? std::string(".") : loadDir.substr(0, s)) + "/restart_files";
  }

  // Array2D::loadtxt (Array.h:2644-2667): nX rows of nY values (the save2file
  // layout) into the reference layout x + XsizeFull * y.
  static bool readTxt2D(const std::string& path, const clProblem& p, std::vector<double>& v) {
    std::ifstream in(path);
    if (!in) return false;
    for (int i = 0; i < p.nX; i++)
      for (int j = 0; j < p.nY; j++)
        if (!(in >> v[size_t(i) + size_t(p.XsizeFull) * j])) return false;
    return true;
  }

  // a .bin file of n doubles (Array.h's binary layout)
  static bool readBin(const std::string& path, size_t n, std::vector<double>& v) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    v.assign(n, 0.);
    const size_t got = std::fread(v.data(), sizeof(double), n, f);
    std::fclose(f);
    return got == n;
  }
};

// clVariablesFD (clVariablesFD.h/.cpp): the temperature solve after the
// k-omega step.  Its keys are in the thermal document (ParamsName 3);
// without that document the solve is off.  Update_T_Coeffs_Implicit reads
// the eddy viscosity, so the reference's temperature kernels need k-omega:
// the solve on with k-omega off is an error.
class clVariablesFD {
 public:
  bool thermalSolverFlag = false, calcNuFlag = true, saveMacroStart = false, calcSSTempFlag = true;
  bool chtCorrectionFlag = true;
  // ConstDef.h:59-63, 232-249, 299-304
  double PrNum = 0.708, PrTurbNum = 0.9, kSoot = 0.057, kAir = 0.042, rhoSoot = 35.,
         cpSoot = 866.8;
  double T_Actual_Max = 400. + 273., T_Actual_Min = 90. + 273., ROE0 = 0.5, ROE_INX = 2.0;
  double tempMaxRelTol = 1e-4, tempMaxAbsTol = 1e-7;
  int tempMaxIters = 1000;
  double nuCutoffRadius = 5.;
  int maxOutputLinesNu = 10;
  double Alpha_fluid = 0., Alpha_foul = 0.;
  bool tempLoaded = false;
  double kSootLattice = 0., kAirLattice = 0.;  // KSOOT, KAIR
  TimeRows Nu, NuMean;
  kofm_ctx* ctx = nullptr;

  bool nuOn() const { return thermalSolverFlag && calcNuFlag; }

  static bool knownKey(const std::string& k) {
    static const char* const keys[] = {
        "ParamsName", "Dummy", "Restart Run", "Thermal Solver", "Calculate Nu",
        "Save Macros On Start", "Solve Steady Temp", "Use CHT Correction", "Pr Number", "K Soot",
        "K air", "Rho Soot", "Cp Soot", "Tinlet Actual", "Twalls Actual", "Tinitial",
        "Tinlet Sim", "Temp Max Rel Tol", "Temp Max Abs Tol", "Temp Max Iterations",
        "Nu Cutoff Radius", "Max Nu Output Lines"};
    for (const char* s : keys)
      if (k == s) return true;
    return false;
  }

  // clVariablesFD::loadParams (clVariablesFD.cpp:445-497).  "Pr Number" sets
  // both Pr and the turbulent Pr (the reference reads PrTurbNum from the same
  // key, :456), which is kept.  "Calculate Nu" (default on) adds the
  // Nusselt record, which needs the wavy section of the system document, as
  // in the reference (iniNuCoeffs, :364-366); "Solve Steady Temp" is read and, as in the reference
  // (whose only call of solveSSThermal is commented out,
  // clVariablesLB.cpp:237), changes nothing; "Use CHT Correction" only adds an
  // argument that the active kernel does not use.
  void loadParams(const clProblem& p, const clVariablesLB& vlb) {
    const RunParams& rp = p.params;
    const int d = RunParams::kThermal;
    bool present = false;
    for (const std::string& k : rp.keys(d)) {
      present = true;
      if (!knownKey(k))
        throw std::runtime_error("unsupported thermal parameter \"" + k + "\" in RunParam.yaml");
    }
    thermalSolverFlag = present && rp.getBool(d, "Thermal Solver", true);
    calcNuFlag = rp.getBool(d, "Calculate Nu", true);
    saveMacroStart = rp.getBool(d, "Save Macros On Start", false);
    calcSSTempFlag = rp.getBool(d, "Solve Steady Temp", true);
    chtCorrectionFlag = rp.getBool(d, "Use CHT Correction", true);
    PrNum = rp.getDouble(d, "Pr Number", 0.708);
    PrTurbNum = rp.getDouble(d, "Pr Number", 0.9);
    Alpha_fluid = vlb.MuVal / PrNum;
    kSoot = rp.getDouble(d, "K Soot", 0.057);
    kAir = rp.getDouble(d, "K air", 0.042);
    rhoSoot = rp.getDouble(d, "Rho Soot", 35.);
    cpSoot = rp.getDouble(d, "Cp Soot", 866.8);
    // (kSoot / rhoSoot / cpSoot) DELTA_L^2 / DELTA_T with DELTA_T =
    // nuAir DELTA_L^2 / MuVal (clVariablesLB.cpp:608)
    Alpha_foul = (kSoot / rhoSoot / cpSoot) * (vlb.MuVal / vlb.nuAir);
    T_Actual_Max = rp.getDouble(d, "Tinlet Actual", 400. + 273.);
    T_Actual_Min = rp.getDouble(d, "Twalls Actual", 90. + 273.);
    ROE0 = rp.getDouble(d, "Tinitial", 0.5);
    ROE_INX = rp.getDouble(d, "Tinlet Sim", 2.0);
    tempMaxRelTol = rp.getDouble(d, "Temp Max Rel Tol", 1e-4);
    tempMaxAbsTol = rp.getDouble(d, "Temp Max Abs Tol", 1e-7);
    tempMaxIters = rp.getInt(d, "Temp Max Iterations", 1000);
    nuCutoffRadius = rp.getDouble(d, "Nu Cutoff Radius", 5.);
    maxOutputLinesNu = rp.getInt(d, "Max Nu Output Lines", 10);
    if (thermalSolverFlag && !vlb.kOmegaClass.kOmegaSolverFlag)
      throw std::runtime_error("the temperature solve needs k-omega (\"Use Turb Model\"): its rows "
                               "take the eddy viscosity");
    // KSOOT, KAIR = k * DELTA_F / DELTA_T (clVariablesTR.cpp:1912-1913) with
    // DELTA_L = 2 R / PIPE_DIAMETER_REAL (clProblem.cpp:484, ConstDef.h:58)
    // and DELTA_T, DELTA_M, DELTA_F of clVariablesLB.cpp:608-610
    const double dL = (2. * p.Pipe_radius) / 2.e-3;
    const double dT = vlb.nuAir * dL * dL / vlb.MuVal;
    const double dM = vlb.RhoVal * dL * dL * dL / vlb.rhoAir;
    const double dF = dM * dL / dT / dT;
    kSootLattice = kSoot * dF / dT;
    kAirLattice = kAir * dF / dT;
  }

  // clVariablesFD::ini (clVariablesFD.cpp:230-283) with testRestartRun
  // (:790-806): load/temp.bin, when present, replaces the initial field.
  void ini(const clProblem& p, kofm_ctx* c, const std::string& loadDir = "load") {
    ctx = c;
    if (!thermalSolverFlag) return;
    kofm_thermal_params q{};
    q.alpha_fluid = Alpha_fluid;
    q.alpha_foul = Alpha_foul;
    q.pr_turb = PrTurbNum;
    q.t_in = ROE_INX;
    q.t_initial = ROE0;
    q.dt = p.dTtfd;
    q.rel_tol = tempMaxRelTol;
    q.abs_tol = tempMaxAbsTol;
    q.max_iters = tempMaxIters;
    check(ctx, kofm_thermal_init(ctx, &q), "clVariablesFD::ini");
    const size_t D = size_t(p.XsizeFull) * p.nY;
    std::vector<double> T;
    tempLoaded = clVariablesLB::readBin(loadDir + "/temp.bin", D, T);
    if (tempLoaded) check(ctx, kofm_set_temperature(ctx, T.data()), "load temp.bin");
  }

  // clVariablesFD::iniNuCoeffs (clVariablesFD.cpp:362-442), called after
  // ini as LBLS.cpp:93-96 calls it; the TimeData files Nu.txt and
  // NuMean.txt start empty unless the run is a restart
  void iniNuCoeffs(const clProblem& p, bool restart, const std::string& outDir) {
    if (!nuOn()) return;
    kofm_nusselt_params q{};
    q.height = p.Channel_Height;
    q.wavy_start = p.xWavyStart;
    q.wavy_period = p.wavyPeriodLen;
    q.wavy_periods = p.numWavyPeriods;
    q.k_soot = kSootLattice;
    q.k_air = kAirLattice;
    check(ctx, kofm_nusselt_init(ctx, &q), "clVariablesFD::iniNuCoeffs");
    Nu.ini("Nu", outDir, maxOutputLinesNu, restart);
    NuMean.ini("NuMean", outDir, maxOutputLinesNu, restart);
  }

  // clVariablesFD::updateTimeData (clVariablesFD.cpp:844-869): one row of
  // each record at this time
  void updateTimeData(unsigned time) {
    if (!nuOn()) return;
    int32_t n = 0, locs = 0;
    check(ctx, kofm_nusselt_sizes(ctx, &n, &locs), "Nu sizes");
    std::vector<double> nu(5 * size_t(n)), mean(2 * size_t(locs));
    check(ctx, kofm_nusselt(ctx, nu.data(), mean.data(), nullptr, nullptr), "FD_calc_Nu");
    Nu.add(time, std::move(nu));
    NuMean.add(time, std::move(mean));
  }

  // clVariablesFD::saveTimeData (clVariablesFD.cpp:562-570)
  void saveTimeData() {
    if (!nuOn()) return;
    Nu.appendData();
    NuMean.appendData();
  }

  // clVariablesFD::Solve (clVariablesFD.h:343-358)
  void Solve() {
    if (thermalSolverFlag) check(ctx, kofm_thermal_step(ctx, 1), "clVariablesFD::Solve");
  }
};

// shearStress (shearStress.h:189-290): node and boundary-link wall shear.
class shearStress {
 public:
  void calculateShear(kofm_ctx* ctx) { check(ctx, kofm_wall_shear(ctx), "shear"); }
  std::vector<double> blTau(kofm_ctx* ctx) {
    kofm_geom_counts c;
    check(ctx, kofm_geometry_counts(ctx, &c), "counts");
    std::vector<double> t(size_t(std::max(c.nBLshear, 1)));
    check(ctx, kofm_download(ctx, KOFM_FIELD_TAU_BL, t.data()), "shear download");
    t.resize(size_t(c.nBLshear));
    return t;
  }
  // TimeData row "time\ttau_0\ttau_1...\n" in ShearStress.txt
  void updateTimeData(kofm_ctx* ctx, unsigned time, const std::string& dir) {
    std::ofstream out(dir + "/ShearStress.txt", std::ios::app);
    out << time;
    for (double v : blTau(ctx)) out << "\t" << v;
    out << "\n";
  }

  // The tracer's defines TR_shear_removal takes (clVariablesTR.cpp:
  // 1328-1373, 1900-1942) from the particle document: the surface-type
  // threshold, the re-release timer and the particle location's row length
  // (the particle domain floor(X Min Val) to floor(X Stop Pos) rounded up to
  // WORKGROUPSIZE_TR = 64).  The per-size arrays are particleProperties' and
  // stay with the tracer.
  static kofm_removal_params removalParams(const clProblem& p) {
    const RunParams& rp = p.params;
    kofm_removal_params q{};
    q.dt_wall = double(p.trSteps_wall);
    q.foul_size_switch = rp.getDouble(RunParams::kParticles, "Foul Size Switch Soot", 0.5);
    // TIME_BEFORE_RERELEASE = 200000 / LB_STEPS_PER_TR_WALL (ConstDef.h:150, 290)
    const int before = rp.getInt(RunParams::kParticles, "Time Before Re-release", 200000 / 3);
    q.timer_start = std::min(before, INT_MAX - 1);
    const double release = rp.getDouble(RunParams::kParticles, "X Release Pos", 90.);
    const double xMin = rp.getDouble(RunParams::kParticles, "X Min Val", release - 1.25);
    const double xStop = rp.getDouble(RunParams::kParticles, "X Stop Pos", 1900.);
    const int width = int(std::floor(xStop)) - int(std::floor(xMin));
    q.loc_stride = (width + 63) / 64 * 64;
    return q;
  }

  // shearStress::shearRemoval with updateParRemArgs' range (shearStress.cpp:
  // 327-335): the particles sorted as temporarily deposited,
  // [Ploc(1).x, Ploc(1).y), after calculateShear on the same stream.
  void shearRemoval(kofm_ctx* ctx, const kofm_removal_params& q, const kofm_particles& P,
                    int64_t first, int64_t end, uint32_t* blDep) {
    check(ctx, kofm_shear_removal(ctx, &q, &P, first, end - first, blDep), "TR_shear_removal");
  }
};

}  // namespace ref
}  // namespace kofm
