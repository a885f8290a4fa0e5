// gpue_host.cpp -- GPUE's host control flow over the MI355X step.
//
// Everything here is host C++: parameters, field construction on the host
// side, file formats and the step loop's I/O cadence.  All device work goes
// through the C ABI (include/gpue_amd.h).
#include "gpue_host.h"

#include <getopt.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

namespace gpue_host {

static const double PI = 3.141592653589793;      // include/constants.h:19
static const double HBAR_SI = 1.05457148e-34;    // include/constants.h:20

static void check(gpue_ctx* c, gpue_status s, const char* what) {
  if (s != GPUE_OK) {
    std::string msg = std::string(what) + " failed (status " + std::to_string((int)s) + ")";
    if (c) msg += ": " + std::string(gpue_last_error(c));
    throw std::runtime_error(msg);
  }
}

// ---------------------------------------------------------------------------
// Grid
// ---------------------------------------------------------------------------
int Grid::ival(const std::string& k) const {
  auto it = ints_.find(k);
  if (it == ints_.end()) throw std::runtime_error("missing int parameter " + k);
  return it->second;
}
double Grid::dval(const std::string& k) const {
  auto it = doubles_.find(k);
  if (it == doubles_.end()) throw std::runtime_error("missing double parameter " + k);
  return it->second;
}
bool Grid::bval(const std::string& k) const {
  auto it = bools_.find(k);
  if (it == bools_.end()) throw std::runtime_error("missing bool parameter " + k);
  return it->second;
}
std::string Grid::sval(const std::string& k) const {
  auto it = strings_.find(k);
  if (it == strings_.end()) throw std::runtime_error("missing string parameter " + k);
  return it->second;
}

// ---------------------------------------------------------------------------
// CLI (src/parser.cu:7-490)
// ---------------------------------------------------------------------------
static std::string filecheck(const std::string& f) {
  struct stat st;
  if (stat(f.c_str(), &st) != 0) throw std::runtime_error("file " + f + " not found");
  return f;
}

Grid parse_args(int argc, char** argv) {
  Grid par;
  // defaults: src/parser.cu:14-67
  par.store("xDim", 256);
  par.store("yDim", 256);
  par.store("zDim", 256);
  par.store("omega", 0.0);
  par.store("gammaY", 1.0);
  par.store("gsteps", 1);
  par.store("esteps", 1);
  par.store("gdt", 1e-4);
  par.store("dt", 1e-4);
  par.store("device", 0);
  par.store("atoms", 1);
  par.store("read_wfc", false);
  par.store("printSteps", 100);
  par.store("winding", 0.0);
  par.store("corotating", false);
  par.store("gpe", false);
  par.store("omegaZ", 6.283);
  par.store("interaction", 1.0);
  par.store("write_it", false);
  par.store("x0_shift", 0.0);
  par.store("y0_shift", 0.0);
  par.store("z0_shift", 0.0);
  par.store("omegaX", 6.283);
  par.store("omegaY", 6.283);
  par.store("data_dir", std::string("data/"));
  par.store("dimnum", 2);
  par.store("write_file", true);
  par.store("fudge", 0.0);
  par.store("mask_2d", 0.0);
  par.store("box_size", -0.01);
  par.store("energy_calc", false);
  par.store("energy_calc_steps", 0);
  par.store("energy_calc_threshold", -1.0);
  par.store("use_param_file", false);
  par.store("param_file", std::string("param.cfg"));
  par.store("thresh_const", 1.0);
  // extensions beyond GPUE's flags: explicit hbar/mass/g for oscillator units, the
  // working precision, and the imaginary->real hand-off policy
  par.store("hbar", HBAR_SI);
  par.store("mass", 1.4431607e-25);   // src/init.cu:98
  par.store("g_direct", std::nan(""));
  par.store("precision", 64);
  par.store("stale_handoff", true);
  par.store("ramp", false);           // src/parser.cu:45-46
  par.store("ramp_type", 1);
  optind = 1;
  int opt;
  // GPUE's short flags of the step (src/parser.cu:71-72) plus -B/-M/-N/-o, letters GPUE
  // leaves unused; the working precision is a long option because every other letter
  // is taken (GPUE's -P is the laser power of its optical-lattice experiments).
  static const struct option longopts[] = {{"precision", required_argument, nullptr, 1000},
                                           {nullptr, 0, nullptr, 0}};
  while ((opt = getopt_long(argc, argv, "b:d:C:x:y:w:m:G:g:e:T:t:n:p:rQ:L:E::lsi:X:Y:WU:V:z:A:v:Z:fc:I:j:B:M:N:oR:",
                            longopts, nullptr)) != -1) {
    switch (opt) {
      case 'x': par.store("xDim", std::atoi(optarg)); break;
      case 'y': par.store("yDim", std::atoi(optarg)); break;
      case 'z': par.store("zDim", std::atoi(optarg)); break;
      case 'b': par.store("box_size", std::atof(optarg)); break;
      case 'j': par.store("thresh_const", std::atof(optarg)); break;
      case 'w': par.store("omega", std::atof(optarg)); break;
      case 'm': par.store("mask_2d", std::atof(optarg)); break;
      case 'G': par.store("gammaY", std::atof(optarg)); break;
      case 'g': par.store("gsteps", (int)std::atof(optarg)); break;
      case 'e': par.store("esteps", (int)std::atof(optarg)); break;
      case 'T': par.store("gdt", std::atof(optarg)); break;
      case 't': par.store("dt", std::atof(optarg)); break;
      case 'C': par.store("device", std::atoi(optarg)); break;
      case 'n': par.store("atoms", (int)std::atof(optarg)); break;
      case 'I':
        par.store("param_file", std::string(optarg));
        par.store("use_param_file", true);
        break;
      case 'r': par.store("read_wfc", true); break;
      case 'p': par.store("printSteps", std::atoi(optarg)); break;
      case 'L': par.store("winding", std::atof(optarg)); break;
      case 'l': par.store("corotating", true); break;
      case 'E':   // -E threshold steps (src/parser.cu:222-239)
        if (!(optind >= argc || argv[optind][0] == '-')) {
          par.store("energy_calc_threshold", std::atof(argv[optind]));
          par.store("energy_calc_steps", optind + 1 < argc ? std::atoi(argv[optind + 1]) : 0);
          optind++;
        }
        par.store("energy_calc", true);
        break;
      case 'f': par.store("write_file", false); break;
      case 's': par.store("gpe", true); break;
      case 'Z': par.store("omegaZ", std::atof(optarg)); break;
      case 'i': par.store("interaction", std::atof(optarg)); break;
      case 'X': par.store("omegaX", std::atof(optarg)); break;
      case 'Y': par.store("omegaY", std::atof(optarg)); break;
      case 'W': par.store("write_it", true); break;
      case 'd': par.store("data_dir", std::string(optarg) + "/"); break;
      case 'U': par.store("x0_shift", std::atof(optarg)); break;
      case 'V': par.store("y0_shift", std::atof(optarg)); break;
      case 'Q': par.store("z0_shift", std::atof(optarg)); break;
      case 'v': par.Vfn = optarg; par.Wfcfn = optarg; break;
      case 'c': {
        int d = std::atoi(optarg);
        if (d == 3) {
          par.Vfn = "3d";
          par.Wfcfn = "3d";
          par.store("box_size", 2.5e-5);
        }
        par.store("dimnum", d);
        break;
      }
      case 'A': par.Afn = optarg; break;
      case 'B': par.store("hbar", std::atof(optarg)); break;
      case 'M': par.store("mass", std::atof(optarg)); break;
      case 'N': par.store("g_direct", std::atof(optarg)); break;
      case 1000: {
        const int prec = std::atoi(optarg);
        if (prec != 64 && prec != 32) throw std::runtime_error("--precision must be 64 or 32");
        par.store("precision", prec);
        break;
      }
      case 'o': par.store("stale_handoff", false); break;
      case 'R':   // src/parser.cu:188-195
        std::printf("Ramping omega with imaginary time evolution\n");
        par.store("ramp", true);
        par.store("ramp_type", std::atoi(optarg));
        break;
      default:
        if (optopt == 0 && optind > 0 && optind <= argc)   // an unknown long option
          throw std::runtime_error(std::string("unsupported option ") + argv[optind - 1]);
        throw std::runtime_error(std::string("unsupported option -") + (char)optopt +
                                 " (analysis/experiment flags of GPUE are outside the step)");
    }
  }
  const std::string data_dir = par.sval("data_dir");
  struct stat st;
  if (stat(data_dir.c_str(), &st) == -1) mkdir(data_dir.c_str(), 0700);
  const int dimnum = par.ival("dimnum");
  if (dimnum < 3) par.store("zDim", 1);
  if (dimnum < 2) par.store("yDim", 1);
  if (par.bval("use_param_file")) par.store("param_file", filecheck(data_dir + par.sval("param_file")));
  if (par.bval("read_wfc")) {
    par.store("infile", filecheck(data_dir + "wfc_load"));
    par.store("infilei", filecheck(data_dir + "wfci_load"));
  }
  if (par.Afn == "file") {
    par.store("Axfile", filecheck(data_dir + "Axgauge"));
    par.store("Ayfile", filecheck(data_dir + "Aygauge"));
    if (dimnum == 3) par.store("Azfile", filecheck(data_dir + "Azgauge"));
  }
  return par;
}

// ---------------------------------------------------------------------------
// equation file (src/dynamic.cu:187-222)
// ---------------------------------------------------------------------------
void parse_param_file(Grid& par, const std::string& path) {
  std::ifstream file(path);
  if (!file.is_open()) throw std::runtime_error("cannot open " + path);
  std::string line, eqn, key;
  auto flush = [&]() {
    if (key.empty()) return;
    par.eqns[key] = eqn;
    par.eqn_order.push_back(key);
  };
  while (std::getline(file, line)) {
    line.erase(std::remove_if(line.begin(), line.end(), [](unsigned char ch) { return std::isspace(ch); }), line.end());
    size_t it = line.find("=");
    if (it == std::string::npos) {
      eqn += line;
    } else {
      if (!key.empty()) {
        if (key == "V") eqn += "+0.5*mass*(Ax*Ax+Ay*Ay+Az*Az)";
        flush();
      }
      key = line.substr(0, it);
      eqn = line.substr(it + 1);
    }
  }
  flush();
}

// ---------------------------------------------------------------------------
// FileIO (src/fileIO.cu)
// ---------------------------------------------------------------------------
namespace FileIO {
std::vector<double> readIn(const std::string& fileR, const std::string& fileI, int gSize) {
  std::vector<double> il(2 * (size_t)gSize, 0.0);
  for (int part = 0; part < 2; ++part) {
    FILE* f = std::fopen((part == 0 ? fileR : fileI).c_str(), "r");
    if (!f) throw std::runtime_error("cannot open wavefunction file");
    double v;
    int i = 0;
    while (i < gSize && std::fscanf(f, "%lE", &v) > 0) il[2 * (size_t)i++ + part] = v;
    std::fclose(f);
  }
  return il;
}

void writeOut(const std::string& file, const double* il, int length, int step) {
  for (int part = 0; part < 2; ++part) {
    std::string name = file + (part ? "i_" : "_") + std::to_string(step);
    FILE* f = std::fopen(name.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + name);
    for (int i = 0; i < length; ++i) std::fprintf(f, "%.16e\n", il[2 * (size_t)i + part]);
    std::fclose(f);
  }
}

void writeOutDouble(const std::string& file, const double* data, int length, int step) {
  std::string name = file + "_" + std::to_string(step);
  FILE* f = std::fopen(name.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + name);
  for (int i = 0; i < length; ++i) std::fprintf(f, "%.16e\n", data[i]);
  std::fclose(f);
}

void writeOutInt(const std::string& file, const int* data, int length, int step) {
  std::string name = file + "_" + std::to_string(step);
  FILE* f = std::fopen(name.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + name);
  for (int i = 0; i < length; ++i) std::fprintf(f, "%d\n", data[i]);
  std::fclose(f);
}

void writeOutVortex(const std::string& file, const std::vector<Vortex>& v, int step) {
  std::string name = file + "_" + std::to_string(step);
  FILE* f = std::fopen(name.c_str(), "w");
  if (!f) throw std::runtime_error("cannot write " + name);
  std::fprintf(f, "#UID,X,Xd,Y,Yd,WINDING,isOn\n");
  for (const Vortex& q : v) std::fprintf(f, "%d,%e,%d,%e,%d\n", q.x, q.xd, q.y, q.yd, q.wind);
  std::fclose(f);
}

// Grid::write (src/ds.cu:269-288): [Params] then key=value for doubles and ints
void writeOutParam(const Grid& par, const std::string& file) {
  std::ofstream out(file);
  out << "[Params]" << "\n";
  for (auto& kv : par.doubles()) out << kv.first << "=" << kv.second << '\n';
  for (auto& kv : par.ints()) out << kv.first << "=" << kv.second << '\n';
}
}  // namespace FileIO

// ---------------------------------------------------------------------------
// Tracker (src/tracker.cu:212-345)
// ---------------------------------------------------------------------------
namespace Tracker {
std::vector<Vortex> vortPos(const int* marker, int xDim) {
  std::vector<Vortex> v;
  for (int i = 0; i < xDim; ++i)
    for (int j = 0; j < xDim; ++j)
      if (std::abs(marker[i * xDim + j]) >= 1) v.push_back({i, j, 0.0, 0.0, marker[i * xDim + j]});
  return v;
}

// least-squares plane through the plaquette corners (include/tracker.h:46-48)
void lsFit(std::vector<Vortex>& v, const double* il, int xDim) {
  static const double lsq[3][4] = {{-0.5, 0.5, -0.5, 0.5}, {-0.5, -0.5, 0.5, 0.5}, {0.75, 0.25, 0.25, -0.25}};
  for (Vortex& q : v) {
    const long idx[4] = {(long)q.x * xDim + q.y, (long)(q.x + 1) * xDim + q.y, (long)q.x * xDim + (q.y + 1),
                         (long)(q.x + 1) * xDim + (q.y + 1)};
    double rx[3], ry[3];
    for (int jj = 0; jj < 3; ++jj) {
      rx[jj] = lsq[jj][0] * il[2 * idx[0]] + lsq[jj][1] * il[2 * idx[1]] + lsq[jj][2] * il[2 * idx[2]] +
               lsq[jj][3] * il[2 * idx[3]];
      ry[jj] = lsq[jj][0] * il[2 * idx[0] + 1] + lsq[jj][1] * il[2 * idx[1] + 1] + lsq[jj][2] * il[2 * idx[2] + 1] +
               lsq[jj][3] * il[2 * idx[3] + 1];
    }
    const double det = 1.0 / (rx[0] * ry[1] - ry[0] * rx[1]);
    const double Rx = det * (ry[1] * rx[2] - ry[0] * ry[2]);
    const double Ry = det * (-rx[1] * rx[2] + rx[0] * ry[2]);
    q.xd = q.x - Rx;
    q.yd = q.y - Ry;
  }
}

// nearest-neighbour re-ordering against the previous list (src/tracker.cu:234-256)
void vortArrange(std::vector<Vortex>& c, const std::vector<Vortex>& p) {
  for (size_t i = 0; i < c.size() && i < p.size(); ++i) {
    int dist = 0x7FFFFFFF;
    size_t index = i;
    for (size_t j = i; j < c.size(); ++j) {
      int dist_t = (int)((p[i].xd - c[j].xd) * (p[i].xd - c[j].xd) + (p[i].yd - c[j].yd) * (p[i].yd - c[j].yd));
      if (dist > dist_t) {
        dist = dist_t;
        index = j;
      }
    }
    std::swap(c[index], c[i]);
  }
}
}  // namespace Tracker

// ---------------------------------------------------------------------------
// init (src/init.cu:40-356)
// ---------------------------------------------------------------------------
gpue_ctx* init(Grid& par) {
  const int dimnum = par.ival("dimnum");
  const int xDim = par.ival("xDim"), yDim = par.ival("yDim"), zDim = par.ival("zDim");
  const int N = par.ival("atoms");
  const double hbar = par.dval("hbar"), mass = par.dval("mass");
  const double omegaX = par.dval("omegaX"), omegaY = par.dval("omegaY"), omegaZ = par.dval("omegaZ");
  const double a_s = 4.76e-9;   // src/init.cu:100
  par.store("a_s", a_s);
  const double a0x = std::sqrt(hbar / (2 * mass * omegaX));
  const double a0y = std::sqrt(hbar / (2 * mass * omegaY));
  const double a0z = std::sqrt(hbar / (2 * mass * omegaZ));
  par.store("a0x", a0x);
  par.store("a0y", a0y);
  par.store("a0z", a0z);
  double gDenConst = N * 4 * hbar * hbar * PI * (a_s / mass);   // src/init.cu:114-117
  if (dimnum == 2) gDenConst *= std::sqrt(mass * (omegaZ) / (2 * PI * hbar));
  par.store("gDenConst", gDenConst);
  const double Rxy = std::pow(15, 0.2) * std::pow(N * a_s * std::sqrt(mass * omegaZ / hbar), 0.2);
  par.store("Rxy", Rxy);
  const double box = par.dval("box_size");
  const double xMax = box > 0 ? box : 6 * Rxy * a0x;
  const double yMax = box > 0 ? box : 6 * Rxy * a0y;
  const double zMax = box > 0 ? box : 6 * Rxy * a0z;
  par.store("xMax", xMax);
  par.store("yMax", yMax);
  par.store("zMax", zMax);

  gpue_params p;
  gpue_params_default(&p);
  p.dimnum = dimnum;
  p.xDim = xDim;
  p.yDim = yDim;
  p.zDim = zDim;
  p.xMax = xMax;
  p.yMax = yMax;
  p.zMax = zMax;
  p.hbar = hbar;
  p.mass = mass;
  const double gd = par.dval("g_direct");
  p.g = std::isnan(gd) ? par.dval("interaction") * gDenConst : gd;
  p.omega = par.dval("omega");
  p.omegaX = omegaX;
  p.omegaY = omegaY;
  p.omegaZ = omegaZ;
  p.gammaY = par.dval("gammaY");
  p.x0_shift = par.dval("x0_shift");
  p.y0_shift = par.dval("y0_shift");
  p.z0_shift = par.dval("z0_shift");
  p.gdt = par.dval("gdt");
  p.dt = par.dval("dt");
  p.gpe = par.bval("gpe") ? 1 : 0;
  p.corotating = par.bval("corotating") ? 1 : 0;
  p.precision = par.ival("precision");
  p.device = par.ival("device");
  if (par.Afn == "rotation") p.a_kind = GPUE_A_ROTATION;
  else if (par.Afn == "constant" || par.Afn == "file") p.a_kind = GPUE_A_CONSTANT;
  else throw std::runtime_error("gauge field '" + par.Afn + "' is not part of the step (rotation, constant, file)");
  if (par.Vfn != "2d" && par.Vfn != "3d") throw std::runtime_error("potential '" + par.Vfn + "' is not supported");

  gpue_ctx* ctx = nullptr;
  check(nullptr, gpue_create(&p, &ctx), "gpue_create");
  double dxyz[3];
  {
    std::vector<double> x(xDim), y(yDim), z(zDim);
    check(ctx, gpue_get_grid(ctx, 0, x.data()), "get_grid");
    check(ctx, gpue_get_grid(ctx, 1, y.data()), "get_grid");
    check(ctx, gpue_get_grid(ctx, 2, z.data()), "get_grid");
    dxyz[0] = xMax / (xDim >> 1);
    dxyz[1] = dimnum > 1 ? yMax / (yDim >> 1) : 1.0;
    dxyz[2] = dimnum > 2 ? zMax / (zDim >> 1) : 1.0;
    par.store("dx", dxyz[0]);
    par.store("dy", dxyz[1]);
    par.store("dz", dxyz[2]);
    const long gSize = (long)xDim * yDim * zDim;
    par.store("gSize", (int)gSize);
    // gauge fields from files (src/operators.cu:505-520, file_A scales by omega)
    if (par.Afn == "file") {
      const char* keys[3] = {"Axfile", "Ayfile", "Azfile"};
      for (int a = 0; a < dimnum; ++a) {
        std::vector<double> A = FileIO::readIn(par.sval(keys[a]), par.sval(keys[a]), (int)gSize);
        std::vector<double> Ar(gSize);
        for (long i = 0; i < gSize; ++i) Ar[i] = A[2 * i] * par.dval("omega");
        check(ctx, gpue_set_field(ctx, GPUE_FIELD_AX + a, Ar.data()), "set_field A");
      }
    }
    // equation fields (-I): Ax/Ay/Az first, then V, K (GPUE builds A before V)
    if (par.bval("use_param_file")) {
      parse_param_file(par, par.sval("param_file"));
      std::vector<std::string> names;
      std::vector<double> vals;
      for (auto& kv : par.doubles()) {
        names.push_back(kv.first);
        vals.push_back(kv.second);
      }
      // simple key = number lines become Grid doubles usable by later equations
      for (auto& key : par.eqn_order) {
        const std::string& e = par.eqns[key];
        if (!e.empty() && e.find_first_not_of("0123456789.") == std::string::npos) {
          par.store(key, std::atof(e.c_str()));
          names.push_back(key);
          vals.push_back(std::atof(e.c_str()));
        }
      }
      std::vector<const char*> cn;
      for (auto& n : names) cn.push_back(n.c_str());
      const char* order[5] = {"Ax", "Ay", "Az", "V", "K"};
      const int which[5] = {GPUE_FIELD_AX, GPUE_FIELD_AY, GPUE_FIELD_AZ, GPUE_FIELD_V, GPUE_FIELD_K};
      for (int k = 0; k < 5; ++k) {
        auto it = par.eqns.find(order[k]);
        if (it == par.eqns.end()) continue;
        std::string e = it->second;
        // references to A fields in V are expanded in place (GPUE inlines earlier equations)
        for (const char* an : {"Ax", "Ay", "Az"}) {
          std::string rep = par.eqns.count(an) ? "(" + par.eqns[an] + ")" : std::string("0");
          size_t pos;
          while ((pos = e.find(an)) != std::string::npos) e.replace(pos, 2, rep);
        }
        check(ctx, gpue_set_field_eqn(ctx, which[k], e.c_str(), cn.data(), vals.data(), (int)cn.size()),
              "set_field_eqn");
      }
    }
    // psi0: kstd_wfc (src/operators.cu:1031-1050) or wfc_load (src/init.cu:525-539)
    if (par.bval("read_wfc")) {
      par.wfc = FileIO::readIn(par.sval("infile"), par.sval("infilei"), (int)gSize);
    } else {
      par.wfc.assign(2 * gSize, 0.0);
      const double winding = par.dval("winding");
      const double w0 = Rxy * Rxy * a0x * a0x, w1 = Rxy * Rxy * a0y * a0y, w2 = Rxy * Rxy * (dimnum == 3 ? a0z * a0z : 1.0);
      double sum = 0.0;
      for (long gid = 0; gid < gSize; ++gid) {
        long xi = gid % xDim, yi = (gid / xDim) % yDim, zi = gid / ((long)xDim * yDim);
        double phi = -std::fmod(winding * std::atan2(y[yi], x[xi]), 2 * PI);
        double amp = std::exp(-(x[xi] * x[xi] / w0 + y[yi] * y[yi] / w1 + z[zi] * z[zi] / w2));
        par.wfc[2 * gid] = amp * std::cos(phi);
        par.wfc[2 * gid + 1] = -amp * std::sin(phi);
        sum += std::sqrt(par.wfc[2 * gid] * par.wfc[2 * gid] + par.wfc[2 * gid + 1] * par.wfc[2 * gid + 1]);
      }
      // GPUE normalises the initial state by sqrt(sum |psi| dV) (src/init.cu:231-233, 301-307)
      sum = std::sqrt(sum * dxyz[0] * dxyz[1] * dxyz[2]);
      for (auto& v : par.wfc) v /= sum;
    }
    if (par.bval("write_file")) {
      const std::string dd = par.sval("data_dir");
      FileIO::writeOutDouble(dd + "x", x.data(), xDim, 0);
      FileIO::writeOutDouble(dd + "y", y.data(), yDim, 0);
      FileIO::writeOutDouble(dd + "z", z.data(), zDim, 0);
      FileIO::writeOut(dd + "WFC", par.wfc.data(), (int)gSize, 0);
    }
  }
  return ctx;
}

void set_variables(Grid& par, gpue_ctx* ctx, int ev_type) {
  (void)ev_type;
  check(ctx, gpue_upload_wfc(ctx, par.wfc.data()), "upload_wfc");
}

// ---------------------------------------------------------------------------
// evolve (src/evolution.cu:260-908)
// ---------------------------------------------------------------------------
void evolve(Grid& par, gpue_ctx* ctx, int numSteps, unsigned gstate, std::string buffer) {
  (void)buffer;
  const std::string data_dir = par.sval("data_dir");
  const int dimnum = par.ival("dimnum");
  const int xDim = par.ival("xDim");
  const long gSize = (long)par.ival("gSize");
  const int printSteps = std::max(1, par.ival("printSteps"));
  const bool energy_calc = par.bval("energy_calc");
  const int ecs = par.ival("energy_calc_steps");
  const int cadence = energy_calc ? (ecs == 0 ? printSteps : ecs) : 0;
  const double thr = par.dval("energy_calc_threshold");
  const bool write_it = par.bval("write_it");
  const bool stale = par.bval("stale_handoff");
  const double mask_2d = par.dval("mask_2d");
  const bool ramp = par.bval("ramp");
  const int ramp_type = par.ival("ramp_type");
  const double omega = par.dval("omega");
  // src/evolution.cu:353,409-428: printed omega_0, the ramp's per-step factor
  auto omega_0 = [&](int i) {
    if (!ramp) return omega * par.dval("omegaX");
    if (ramp_type == 1) return i == 0 ? omega : (double)i / (double)(i + 1);
    return i == 0 ? omega / (double)numSteps : (double)(i + 1) / (double)i;
  };
  check(ctx, gpue_set_ramp(ctx, ramp ? ramp_type : 0, numSteps), "ramp");
  // src/evolution.cu:445-460,705-716
  const std::string fileName = gstate == 0 ? (ramp ? "wfc_0_ramp" : "wfc_0_const") : (ramp ? "wfc_ev_ramp" : "wfc_ev");
  const std::string efile = data_dir + (gstate == 0 ? "energyi.dat" : "energy.dat");
  std::printf("Timestep for %s set as: %E\n", gstate == 0 ? "groundstate solver" : "evolution",
              gstate == 0 ? par.dval("gdt") : par.dval("dt"));
  auto t0 = std::chrono::steady_clock::now();
  std::vector<double> host(2 * gSize);
  std::vector<int> marker((size_t)xDim * xDim);
  std::vector<Vortex> prev;
  std::vector<double> energies(1 + numSteps / std::max(1, cadence));
  int i = 0;
  bool stopped = false;
  while (i < numSteps && !stopped) {
    if (i % printSteps == 0) {
      check(ctx, gpue_download_wfc(ctx, host.data()), "download");
      std::printf("Step: %d    Omega: %lf\n", i, omega_0(i));
      double spent = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      std::printf("Time spent: %lf\n", spent);
      if (gstate == 1 && dimnum == 2) {   // 2D vortex tracking (src/evolution.cu:483-700)
        int found = 0;
        check(ctx, gpue_vortex_marker_2d(ctx, mask_2d, marker.data(), &found), "findVortex");
        std::vector<Vortex> cur = Tracker::vortPos(marker.data(), xDim);
        if (found > 0) {
          Tracker::lsFit(cur, host.data(), xDim);
          if (i != 0) {
            Tracker::vortArrange(cur, prev);
            if (write_it) FileIO::writeOutInt(data_dir + "vLoc_", marker.data(), (int)marker.size(), i);
          }
        }
        if (write_it) FileIO::writeOutVortex(data_dir + "vort_arr", cur, i);
        std::printf("Located %lu vortices\n", (unsigned long)cur.size());
        prev = cur;
      }
      if (write_it) FileIO::writeOut(data_dir + fileName, host.data(), (int)gSize, i);
    }
    // run up to the next print point; with the one-step-stale hand-off of
    // src/evolution.cu:430,906 the host copy must be the state before the last
    // step, so chunks also stop before every step that may end the loop
    int next = std::min(numSteps, (i / printSteps + 1) * printSteps);
    if (stale && gstate == 0) {
      int lim = numSteps - 1;
      if (cadence > 0) lim = std::min(lim, ((i + cadence - 1) / cadence) * cadence);
      if (lim > i) next = std::min(next, lim);
      else {
        check(ctx, gpue_download_wfc(ctx, par.wfc.data()), "download");
        next = i + 1;
      }
    }
    gpue_evolve_result r;
    check(ctx, gpue_evolve(ctx, i, next - i, (int)gstate, cadence, thr, energies.data(), (int)energies.size(), &r),
          "gpue_evolve");
    for (int k = 0; k < r.n_energies; ++k) {
      const int step = ((i + cadence - 1) / cadence + k) * cadence;
      std::printf("Energy[t@%d]=%E\n", step, energies[k]);
      std::ofstream eo;
      if (step == 0) eo.open(efile);
      else eo.open(efile, std::ios::out | std::ios::app);
      eo << energies[k] << '\n';
    }
    if (r.early_stop) {
      std::printf("Stopping early at step %d with energy %E\n", i + r.steps_done - 1, r.energy);
      stopped = true;
    }
    i += r.steps_done;
  }
  if (!(stale && gstate == 0)) check(ctx, gpue_download_wfc(ctx, par.wfc.data()), "download");
}

}  // namespace gpue_host
