// runparams.hpp -- reader/writer for the reference's multi-document
// RunParam.yaml (clProblem.cpp:556-670): documents separated by "---", each
// tagged "ParamsName: n" (1 system, 2 fluid, 3 thermal, 5 level set,
// 6 particles, 7 fouling; clProblem.h:24-30), flat "key: value" entries with
// optional !!int / !!float tags, True/False booleans and [a, b] lists.
#pragma once
#include <cstdio>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace kofm {

class RunParams {
 public:
  enum Doc { kSystem = 1, kFluid = 2, kThermal = 3, kLevelSet = 5, kParticles = 6, kFouling = 7 };

  bool load(const std::string& path) {
    std::ifstream in(path);
    if (!in) return false;
    docs_.clear();
    tags_.clear();
    std::map<std::string, std::string> cur, curTags;
    std::string line;
    auto flush = [&]() {
      if (cur.count("ParamsName")) {
        const int doc = std::stoi(cur["ParamsName"]);
        docs_[doc] = cur;
        tags_[doc] = curTags;
      }
      cur.clear();
      curTags.clear();
    };
    while (std::getline(in, line)) {
      if (line.compare(0, 3, "---") == 0) {
        flush();
        continue;
      }
      const size_t hash = line.find('#');
      if (hash == 0) continue;
      const size_t colon = line.find(':');
      if (colon == std::string::npos) continue;
      std::string key = trim(line.substr(0, colon));
      std::string val = trim(line.substr(colon + 1));
      std::string tag;
      if (val.compare(0, 2, "!!") == 0) {  // a type tag, kept for save()
        const size_t sp = val.find(' ');
        tag = val.substr(0, sp);
        val = sp == std::string::npos ? "" : trim(val.substr(sp + 1));
      }
      if (!key.empty()) {
        cur[key] = val;
        if (!tag.empty()) curTags[key] = tag;
      }
    }
    flush();
    return true;
  }

  bool has(int doc, const std::string& key) const {
    auto d = docs_.find(doc);
    return d != docs_.end() && d->second.count(key);
  }

  std::string raw(int doc, const std::string& key) const { return docs_.at(doc).at(key); }

  // Keys of one document (empty if the document is absent).
  std::vector<std::string> keys(int doc) const {
    std::vector<std::string> k;
    auto d = docs_.find(doc);
    if (d != docs_.end())
      for (const auto& kv : d->second) k.push_back(kv.first);
    return k;
  }

  double getDouble(int doc, const std::string& key, double def) const {
    return has(doc, key) ? std::stod(raw(doc, key)) : def;
  }
  int getInt(int doc, const std::string& key, int def) const {
    return has(doc, key) ? int(std::stod(raw(doc, key))) : def;
  }
  bool getBool(int doc, const std::string& key, bool def) const {
    if (!has(doc, key)) return def;
    const std::string v = raw(doc, key);
    return v == "True" || v == "true" || v == "1" || v == "yes";
  }
  std::string getString(int doc, const std::string& key, const std::string& def) const {
    return has(doc, key) ? raw(doc, key) : def;
  }
  // Either of two mutually exclusive keys (clProblem.h:170-190): returns 0 if
  // neither is present, 1 or 2 for the one found; both present is an error.
  int either(int doc, const std::string& k1, const std::string& k2, double& v1, double& v2) const {
    const bool a = has(doc, k1), b = has(doc, k2);
    if (a && b) throw std::runtime_error("Cannot define both " + k1 + " and " + k2);
    if (a) v1 = std::stod(raw(doc, k1));
    if (b) v2 = std::stod(raw(doc, k2));
    return a ? 1 : (b ? 2 : 0);
  }

  // Sets a value with its type tag ("!!int", "!!float" or none).
  void set(int doc, const std::string& key, const std::string& tag, const std::string& val) {
    docs_[doc]["ParamsName"] = std::to_string(doc);
    docs_[doc][key] = val;
    if (tag.empty())
      tags_[doc].erase(key);
    else
      tags_[doc][key] = tag;
  }
  void setDouble(int doc, const std::string& key, double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    set(doc, key, "!!float", buf);
  }
  void setInt(int doc, const std::string& key, int v) { set(doc, key, "!!int", std::to_string(v)); }
  void setBool(int doc, const std::string& key, bool v) { set(doc, key, "", v ? "True" : "False"); }

  // Writes the documents back in the reference's layout (clProblem.cpp:256-405),
  // each value with the type tag it was read or set with.
  bool save(const std::string& path) const {
    std::ofstream out(path);
    if (!out) return false;
    out << "# Parameters for Fouling simulation\n";
    for (const auto& d : docs_) {
      out << "---\nParamsName: " << d.first << "\n";
      auto t = tags_.find(d.first);
      for (const auto& kv : d.second) {
        if (kv.first == "ParamsName") continue;
        out << kv.first << ": ";
        if (t != tags_.end()) {
          auto tk = t->second.find(kv.first);
          if (tk != t->second.end()) out << tk->second << " ";
        }
        out << kv.second << "\n";
      }
      out << "\n";
    }
    return bool(out);
  }

 private:
  std::map<int, std::map<std::string, std::string>> docs_;
  std::map<int, std::map<std::string, std::string>> tags_;
  static std::string trim(const std::string& s) {
    const size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
  }
};

}  // namespace kofm
