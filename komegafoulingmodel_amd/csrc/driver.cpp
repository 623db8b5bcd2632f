// driver.cpp -- kofm_run: the reference's ini/run sequence (LBLS.cpp:65-381)
// for the lattice-Boltzmann step, through the C++ facades.
//
//   kofm_run [load_dir] [out_dir]
//
// reads <load_dir>/RunParam.yaml, lsc0.txt, lsbl.txt (and lbf.bin,
// lbkappa.bin, lbomega.bin for a restart), runs "Ini LB Steps" warm-up steps
// and then steps to "Stop Time", computing the wall shear every
// "LB Steps Per TR Wall" steps, appending the reductions every "Save Avgs"
// interval, the Nusselt record (with the temperature solve) every "Save Nu"
// interval and writing the save2file dumps every "Dump" interval.
#include <algorithm>
#include <cstdio>
#include <exception>
#include <string>

#include "facade.hpp"

// "kofm_run --params <load_dir>" prints the parameters derived from
// RunParam.yaml and exits without touching a GPU.
static int printParams(const std::string& loadDir) {
  using namespace kofm::ref;
  clProblem p;
  p.loadRunParameters(loadDir + "/RunParam.yaml");
  clVariablesLB vlb;
  vlb.loadParams(p);
  std::printf("nX %d\nnY %d\nXsizeFull %d\nPipe_radius %.17g\nChannel_Height %d\n", p.nX, p.nY,
              p.XsizeFull, p.Pipe_radius, p.Channel_Height);
  std::printf("tau0 %.17g\nMuVal %.17g\nRe %.17g\nFval %.17g\nRho %.17g\n", vlb.tau, vlb.MuVal,
              vlb.Re, vlb.Fval, vlb.RhoVal);
  std::printf("UseTurbModel %d\nStopTime %u\ntrStepsWall %u\n",
              vlb.kOmegaClass.kOmegaSolverFlag ? 1 : 0, p.StopTime, p.trSteps_wall);
  clVariablesFD vfd;
  vfd.loadParams(p, vlb);
  std::printf("ThermalSolver %d\nAlphaFluid %.17g\nAlphaFoul %.17g\nPrTurb %.17g\nTin %.17g\n"
              "Tinitial %.17g\ntfdSteps %u\n", vfd.thermalSolverFlag ? 1 : 0, vfd.Alpha_fluid,
              vfd.Alpha_foul, vfd.PrTurbNum, vfd.ROE_INX, vfd.ROE0, p.tfdSteps);
  std::printf("CalcNu %d\nKSootLattice %.17g\nKAirLattice %.17g\nWavyStart %d\nPeriodLength %d\n"
              "NumPeriods %d\nSaveNuTimestep %u\nnextSaveNuTime %u\nnextSaveAvgTime %u\n",
              vfd.nuOn() ? 1 : 0, vfd.kSootLattice, vfd.kAirLattice, p.xWavyStart, p.wavyPeriodLen,
              p.numWavyPeriods, p.saveTimeStepNu, p.nextSaveNuTime, p.nextSaveAvgTime);
  const kofm_removal_params rq = shearStress::removalParams(p);
  std::printf("RemovalDtWall %.17g\nFoulSizeSwitch %.17g\nParTimerStart %d\nParLocStride %d\n",
              rq.dt_wall, rq.foul_size_switch, rq.timer_start, rq.loc_stride);
  std::printf("IniPoiseuille %d\nIniTurbVelocity %d\nPerturbVelField %d\nSaveAvgInfo %d\n"
              "LoadVelText %d\nUseTurbVelBC %d\nGeomPenaltyFactor %.17g\nlbOutMaxLines %d\n",
              vlb.iniPoiseuille, vlb.kOmegaClass.iniTurbVelocity, vlb.kOmegaClass.perturbVelField,
              vlb.saveAvgInfo, vlb.loadVelTxtFlag, vlb.turbVelBCFlag, vlb.geomPenaltyFactor,
              vlb.lbOutSize);
  std::printf("DUplus %.17g\nEpsilonMult %.17g\nBetaMult %.17g\nAlphaMult %.17g\nSigmaMult %.17g\n",
              vlb.kOmegaClass.perturbDUPlus, vlb.kOmegaClass.perturbEpsilon,
              vlb.kOmegaClass.perturbBeta, vlb.kOmegaClass.perturbAlpha,
              vlb.kOmegaClass.perturbSigma);
  return 0;
}

int main(int argc, char** argv) {
  using namespace kofm::ref;
  if (argc > 2 && std::string(argv[1]) == "--params") {
    try {
      return printParams(argv[2]);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "kofm_run: %s\n", e.what());
      return 1;
    }
  }
  const std::string loadDir = argc > 1 ? argv[1] : "load";
  const std::string outDir = argc > 2 ? argv[2] : ".";
  try {
    clProblem p;
    p.loadRunParameters(loadDir + "/RunParam.yaml");
    clVariablesLS vls;
    vls.loadParams(p, loadDir);
    clVariablesLB vlb;
    vlb.loadParams(p);
    clVariablesFD vfd;
    vfd.loadParams(p, vlb);
    vlb.ini(p, vls, loadDir, outDir);
    // "Save Macros On Start" (clVariablesLB.cpp:362-363)
    if (vlb.saveMacroStart && !vlb.restarted) vlb.save2file(outDir);
    vfd.ini(p, vlb.ctx, loadDir);
    vfd.iniNuCoeffs(p, vlb.restarted, outDir);
    // clVariablesFD::ini's "Save Macros On Start" (clVariablesFD.cpp:279-280):
    // the dump then carries the initial temperature too
    if (vfd.thermalSolverFlag && vfd.saveMacroStart && !vlb.restarted) vlb.save2file(outDir);
    std::printf("nX %d nY %d tau0 %.6f nu %.6g Fval %.6g k-omega %s temperature %s Nu %s\n", p.nX,
                p.nY, vlb.tau, vlb.MuVal, vlb.Fval, vlb.kOmegaClass.kOmegaSolverFlag ? "on" : "off",
                vfd.thermalSolverFlag ? "on" : "off", vfd.nuOn() ? "on" : "off");
    if (p.RestartTime == 0 && vlb.tlbmNumIniSteps_LB > 0) vlb.getInitialFlow(p, outDir);
    shearStress ss;
    // a run restarted from a RunParam.yaml without the cadence keys resumes
    // at the next cadence point after its current time
    while (p.nextSaveAvgTime <= p.Time) p.nextSaveAvgTime += p.saveTimeStepAvg;
    while (p.nextSaveNuTime <= p.Time) p.nextSaveNuTime += p.saveTimeStepNu;
    while (p.nextDumpStepTime <= p.Time) p.nextDumpStepTime += p.dumpTimeStep;
    while (p.Time < p.StopTime) {
      unsigned n = std::min(p.StopTime - p.Time, p.trSteps_wall);
      n = std::min(n, p.nextSaveAvgTime - p.Time);
      n = std::min(n, p.nextDumpStepTime - p.Time);
      if (vfd.nuOn()) n = std::min(n, p.nextSaveNuTime - p.Time);
      // the temperature solve follows the LB step every "LB Steps Per T"
      // steps (LBLS.cpp:197-200, 230-233, 315-318)
      if (vfd.thermalSolverFlag) n = std::min(n, p.tfdSteps - p.Time % p.tfdSteps);
      vlb.Solve(int(n));
      p.Time += n;
      if (vfd.thermalSolverFlag && p.Time % p.tfdSteps == 0) vfd.Solve();
      if (p.Time % p.trSteps_wall == 0) ss.calculateShear(vlb.ctx);
      if (p.Time >= p.nextSaveAvgTime) {
        // rows held in memory, written out every "lbOut Max Lines" rows
        vlb.updateTimeData(p.Time);
        ss.updateTimeData(vlb.ctx, p.Time, outDir);
        p.nextSaveAvgTime += p.saveTimeStepAvg;
      }
      // the Nusselt record (clProblem.cpp:110-114)
      if (vfd.nuOn() && p.Time >= p.nextSaveNuTime) {
        vfd.updateTimeData(p.Time);
        p.nextSaveNuTime += p.saveTimeStepNu;
      }
      if (p.Time >= p.nextDumpStepTime) {
        // clProblem::step's dump (clProblem.cpp:130-168): the save2file text
        // dumps and the held time data, RunParam.yaml with the current time
        // and the cadence state; every "Save Bin Step" dumps the restart
        // files and RunParam.yaml go to the load directory, so the next run
        // restarts from them
        p.nextDumpStepTime += p.dumpTimeStep;
        const bool bin = ++p.currentDumpStep >= p.nextDumpBin;
        if (bin) p.nextDumpBin += int(p.saveBinStep);
        vlb.save2file(outDir);
        vlb.saveTimeData();
        vfd.saveTimeData();
        p.saveParameters(outDir);
        if (bin) {
          vlb.saveRestartFiles(loadDir);
          p.saveParameters(loadDir);
        }
      }
    }
    vlb.save2file(outDir);
    vlb.saveTimeData();
    vfd.saveTimeData();
    p.saveParameters(outDir);
    std::printf("done at time %u, Umean %g\n", p.Time, vlb.calcUmean());
  } catch (const std::exception& e) {
    std::fprintf(stderr, "kofm_run: %s\n", e.what());
    return 1;
  }
  return 0;
}
