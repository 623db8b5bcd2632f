// io.cpp -- reference-compatible output files.
#include "io.hpp"

#include <cstdio>
#include <fstream>
#include <sstream>

#include "../../include/kofm.h"
#include "engine.hpp"

namespace kofm {

int writeBin(const std::string& path, const std::vector<double>& a, std::string& err) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    err = "cannot open " + path;
    return KOFM_ERR_ARG;
  }
  size_t n = std::fwrite(a.data(), sizeof(double), a.size(), f);
  std::fclose(f);
  if (n != a.size()) {
    err = "short write to " + path;
    return KOFM_ERR_ARG;
  }
  return 0;
}

int readBin(const std::string& path, std::vector<double>& a, size_t expect, std::string& err) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) {
    err = "cannot open " + path;
    return KOFM_ERR_ARG;
  }
  a.assign(expect, 0.);
  size_t n = std::fread(a.data(), sizeof(double), expect, f);
  std::fclose(f);
  if (n != expect) {
    err = path + " is shorter than the lattice";
    return KOFM_ERR_ARG;
  }
  return 0;
}

int writeTxt2D(const std::string& path, const std::vector<double>& a, int nX, int XF, int nY,
               std::string& err) {
  std::ofstream out(path);
  if (!out) {
    err = "cannot open " + path;
    return KOFM_ERR_ARG;
  }
  for (int i = 0; i < nX; i++) {
    for (int j = 0; j < nY; j++) out << a[i + size_t(XF) * j] << "\t";
    out << "\n";
  }
  return out ? 0 : KOFM_ERR_ARG;
}

int writeTxt2DInt(const std::string& path, const std::vector<int16_t>& a, int nX, int XF, int nY,
                  std::string& err) {
  std::ofstream out(path);
  if (!out) {
    err = "cannot open " + path;
    return KOFM_ERR_ARG;
  }
  for (int i = 0; i < nX; i++) {
    for (int j = 0; j < nY; j++) out << int(a[i + size_t(XF) * j]) << "\t";
    out << "\n";
  }
  return out ? 0 : KOFM_ERR_ARG;
}

int saveFields(Engine& e, int nX, int nY, bool komega, const std::string& dir, std::string& err) {
  const int XF = ((nX + 127) / 128) * 128;
  const size_t D = size_t(XF) * nY;
  std::vector<double> f(9 * D);
  int rc = e.download(KOFM_FIELD_F, f.data(), err);
  if (!rc) rc = writeBin(dir + "/lbf.bin", f, err);
  std::vector<double> a(D);
  struct Item {
    int field;
    const char* txt;
    const char* bin;
  };
  const Item items[] = {{KOFM_FIELD_UX, "lbux", nullptr},
                        {KOFM_FIELD_UY, "lbuy", nullptr},
                        {KOFM_FIELD_RHO, "lbro", nullptr},
                        {KOFM_FIELD_KAPPA, "lbkappa", "lbkappa"},
                        {KOFM_FIELD_OMEGA, "lbomega", "lbomega"}};
  for (const Item& it : items) {
    if (rc) break;
    if (!komega && (it.field == KOFM_FIELD_KAPPA || it.field == KOFM_FIELD_OMEGA)) continue;
    rc = e.download(it.field, a.data(), err);
    if (!rc) rc = writeTxt2D(dir + "/" + it.txt + ".txt", a, nX, XF, nY, err);
    if (!rc && it.bin) rc = writeBin(dir + "/" + it.bin + ".bin", a, err);
  }
  // clVariablesFD::save2file (clVariablesFD.cpp:545-548): the temperature
  // array "temp" in the same text layout
  if (!rc && e.thermalOn()) {
    rc = e.download(KOFM_FIELD_TEMP, a.data(), err);
    if (!rc) rc = writeTxt2D(dir + "/temp.txt", a, nX, XF, nY, err);
  }
  return rc;
}

int saveRestart(Engine& e, int nX, int nY, bool komega, const std::vector<double>& C,
                const std::string& dir, std::string& err) {
  const int XF = ((nX + 127) / 128) * 128;
  const size_t D = size_t(XF) * nY;
  std::vector<double> f(9 * D);
  int rc = e.download(KOFM_FIELD_F, f.data(), err);
  if (!rc) rc = writeBin(dir + "/lbf.bin", f, err);
  if (!rc && komega) {
    std::vector<double> a(D);
    rc = e.download(KOFM_FIELD_KAPPA, a.data(), err);
    if (!rc) rc = writeBin(dir + "/lbkappa.bin", a, err);
    if (!rc) rc = e.download(KOFM_FIELD_OMEGA, a.data(), err);
    if (!rc) rc = writeBin(dir + "/lbomega.bin", a, err);
  }
  // clVariablesFD::saveRestartFiles (clVariablesFD.cpp:550-553): temp.bin,
  // which clVariablesFD::testRestartRun reads back (:797-806)
  if (!rc && e.thermalOn()) {
    std::vector<double> t(D);
    rc = e.download(KOFM_FIELD_TEMP, t.data(), err);
    if (!rc) rc = writeBin(dir + "/temp.bin", t, err);
  }
  if (!rc) rc = writeBin(dir + "/lsc.bin", C, err);
  return rc;
}

int saveDebug(Engine& e, const Geometry& g, bool komega, const std::string& dir, std::string& err) {
  const int nX = g.nX, nY = g.nY, XF = g.XF;
  const size_t D = size_t(g.D);
  int rc = writeTxt2DInt(dir + "/nType.txt", g.nType, nX, XF, nY, err);
  std::vector<double> a(D);
  for (int k = 0; k < 4 && !rc; k++) {
    std::copy(g.dX.begin() + long(k * D), g.dX.begin() + long((k + 1) * D), a.begin());
    rc = writeTxt2D(dir + "/dXarr_" + std::to_string(k) + ".txt", a, nX, XF, nY, err);
  }
  if (!rc && komega) {
    rc = e.download(KOFM_FIELD_NUT, a.data(), err);
    if (!rc) rc = writeTxt2D(dir + "/lbnut.txt", a, nX, XF, nY, err);
    if (!rc) rc = e.download(KOFM_FIELD_WALLD, a.data(), err);
    if (!rc) rc = writeTxt2D(dir + "/wallD.txt", a, nX, XF, nY, err);
  }
  return rc;
}

}  // namespace kofm
