// io.hpp -- reference file formats: raw .bin dumps and save2file text.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "geometry.hpp"

namespace kofm {

class Engine;

// Raw FP64 dump with no header, x padding included (Array.h:1178-1194).
int writeBin(const std::string& path, const std::vector<double>& a, std::string& err);
int readBin(const std::string& path, std::vector<double>& a, size_t expect, std::string& err);

// save2file text for a 2-D field: one row per x < nX, the nY values of
// that column separated by tabs, a trailing tab, newline; default iostream
// formatting (6 significant digits) (Array.h:2528-2546).
int writeTxt2D(const std::string& path, const std::vector<double>& a, int nX, int XF, int nY,
               std::string& err);

// The same for an int16 field (nType), values printed as integers
// (save_txt_from_device_short_to_int, Array.h).
int writeTxt2DInt(const std::string& path, const std::vector<int16_t>& a, int nX, int XF, int nY,
                  std::string& err);

// lbf.bin (+ lbkappa.bin, lbomega.bin) and lbux/lbuy/lbro(/lbkappa/lbomega)
// text dumps into dir.
int saveFields(Engine& e, int nX, int nY, bool komega, const std::string& dir, std::string& err);

// saveRestartFiles of clVariablesLB, kOmega and clVariablesLS
// (clVariablesLB.cpp:752-765, kOmega.cpp:492-496, clVariablesLS.cpp:534-537):
// lbf.bin, lbkappa.bin, lbomega.bin and lsc.bin (the wall points, x y
// interleaved).
int saveRestart(Engine& e, int nX, int nY, bool komega, const std::vector<double>& C,
                const std::string& dir, std::string& err);

// The debug dumps of clVariablesLS::saveDebug and kOmega::saveDebug
// (clVariablesLS.cpp:545-556, kOmega.cpp:422-480): nType.txt (integers),
// dXarr_0..3.txt (one file per plane, Array.h:3193-3220), lbnut.txt and
// wallD.txt.
int saveDebug(Engine& e, const Geometry& g, bool komega, const std::string& dir, std::string& err);

}  // namespace kofm
