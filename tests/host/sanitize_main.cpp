// Host-side checks under AddressSanitizer and UndefinedBehaviorSanitizer: the
// equation parser, its evaluator and postfix lowering (accepted and rejected
// strings), the grid and twiddle builders.  Built and run by
// tests/test_host_sanitizers.py with g++ (no HIP code involved).
#include <cmath>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "../../gpue_amd/csrc/eqn.h"

namespace gpue {
void make_grids(int dimnum, int xDim, int yDim, int zDim, double xMax, double yMax, double zMax,
                std::vector<double>* g, double* d);
void vortex_mask(const double* x, int xDim, double radius, std::vector<unsigned char>& m);
void twiddles(int n, std::vector<double>& re, std::vector<double>& im);
}  // namespace gpue

using namespace gpue;

int main() {
  std::map<std::string, double> d = {{"omega", 0.5}, {"omegaX", 6.283}, {"rad", 10.0}, {"omegaR", 7071.0}};
  std::map<std::string, std::vector<EqnNodeC>> earlier;
  const char* exprs[] = {"(((3*x)+7)+(5-7)+cos(0)*1)+pow(120,2)", "rad*omegaR*x*y*z", "1-2+3", "10-4-3",
                         "8/2*2", "2^3^2", "[2+3]*4", "sqrt(16)+exp(0)", "exp(0-x*x)", "x-y-t",
                         "(x+1)/(y+2)/3", "tan(0.1*x)+pow(y,3)", "[x+y]*[x-y]", "((((x))))",
                         "x^y^x^y^x^y^x^y^x", "jn(1,x)+yn(2,y+2)", "", "(x+1", "x+1)", "foo*2", "k2n(1,2)",
                         "x+", "exp(-x*x)", "((", "))", "[(])", ",", "pow(,)", "sin()", "1..2", "x*y*"};
  int parsed = 0;
  for (const char* e : exprs) {
    std::vector<EqnNodeC> nodes;
    bool td = false;
    std::string err;
    if (eqn_parse(e, d, earlier, nodes, td, err) == 0 && !nodes.empty()) {
      volatile double v = eqn_eval(nodes.data(), (int)nodes.size(), 0.3, -0.7, 1.1, 0.25);
      (void)v;
      EqnProgH prog;
      std::string e2;
      eqn_to_program(nodes.data(), (int)nodes.size(), prog, e2);
      earlier["f" + std::to_string(parsed)] = nodes;
      ++parsed;
    }
  }
  for (int dim = 1; dim <= 3; ++dim)
    for (int n : {16, 48, 512, 4096}) {
      std::vector<double> g[6];
      double dd[3];
      make_grids(dim, n, dim > 1 ? n : 1, dim > 2 ? 16 : 1, 8.0, 6.0, 4.0, g, dd);
      std::vector<double> re, im;
      twiddles(n, re, im);
    }
  std::vector<double> g[6];
  double dd[3];
  make_grids(2, 64, 64, 1, 8.0, 8.0, 1.0, g, dd);
  std::vector<unsigned char> m;
  vortex_mask(g[0].data(), 64, 5.0, m);
  std::printf("parsed %d\n", parsed);
  return 0;
}
