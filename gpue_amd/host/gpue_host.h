// gpue_host.h -- C++ host side of the drop-in: GPUE's parameter store, CLI,
// equation-file reader, file formats and evolve() control flow, driving the
// MI355X step through the C ABI of libgpue_amd.so (include/gpue_amd.h).
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "gpue_amd.h"

namespace gpue_host {

// Grid (include/ds.h:86-197): string-keyed parameter store.  Missing keys throw
// (GPUE asserts, src/ds.cu:191-265) instead of aborting the process.
class Grid {
 public:
  void store(const std::string& k, int v) { ints_[k] = v; }
  void store(const std::string& k, double v) { doubles_[k] = v; }
  void store(const std::string& k, bool v) { bools_[k] = v; }
  void store(const std::string& k, const std::string& v) { strings_[k] = v; }
  void store(const std::string& k, const char* v) { strings_[k] = v; }
  int ival(const std::string& k) const;
  double dval(const std::string& k) const;
  bool bval(const std::string& k) const;
  std::string sval(const std::string& k) const;
  bool is_double(const std::string& k) const { return doubles_.count(k) != 0; }
  const std::map<std::string, double>& doubles() const { return doubles_; }
  const std::map<std::string, int>& ints() const { return ints_; }
  std::string Afn = "rotation", Vfn = "2d", Wfcfn = "2d";
  std::vector<double> wfc;          // host psi, interleaved (re, im)
  std::map<std::string, std::string> eqns;   // equation strings from -I, in file order
  std::vector<std::string> eqn_order;

 private:
  std::map<std::string, int> ints_;
  std::map<std::string, double> doubles_;
  std::map<std::string, bool> bools_;
  std::map<std::string, std::string> strings_;
};

struct Vortex {     // Vtx::Vortex (include/vort.h:39): coords, coordsD, winding
  int x, y;
  double xd, yd;
  int wind;
};

// src/parser.cu:7-490 (flags of the step; analysis/experiment flags are rejected)
Grid parse_args(int argc, char** argv);
// src/dynamic.cu:187-222 (key = expr lines, whitespace removed, continuation lines,
// "+0.5*mass*(Ax*Ax+Ay*Ay+Az*Az)" appended to a V that is not the last entry)
void parse_param_file(Grid& par, const std::string& path);
// src/init.cu:40-356: constants, box, grids, fields and psi0; builds the context
gpue_ctx* init(Grid& par);
// src/init.cu:358-497 hand-off to a phase (uploads the host psi)
void set_variables(Grid& par, gpue_ctx* ctx, int ev_type);
// src/evolution.cu:260-908
void evolve(Grid& par, gpue_ctx* ctx, int numSteps, unsigned gstate, std::string buffer);

namespace FileIO {   // src/fileIO.cu
std::vector<double> readIn(const std::string& fileR, const std::string& fileI, int gSize);
void writeOut(const std::string& file, const double* il, int length, int step);
void writeOutDouble(const std::string& file, const double* data, int length, int step);
void writeOutInt(const std::string& file, const int* data, int length, int step);
void writeOutVortex(const std::string& file, const std::vector<Vortex>& v, int step);
void writeOutParam(const Grid& par, const std::string& file);
}  // namespace FileIO

namespace Tracker {  // src/tracker.cu
std::vector<Vortex> vortPos(const int* marker, int xDim);
void lsFit(std::vector<Vortex>& v, const double* il, int xDim);
void vortArrange(std::vector<Vortex>& cur, const std::vector<Vortex>& prev);
}  // namespace Tracker

}  // namespace gpue_host
