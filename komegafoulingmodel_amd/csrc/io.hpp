// io.hpp -- reference file formats: raw .bin dumps and save2file text.
#pragma once
#include <string>
#include <vector>

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

// lbf.bin (+ lbkappa.bin, lbomega.bin) and lbux/lbuy/lbro(/lbkappa/lbomega)
// text dumps into dir.
int saveFields(Engine& e, int nX, int nY, bool komega, const std::string& dir, std::string& err);

}  // namespace kofm
