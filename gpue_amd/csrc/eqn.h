// eqn.h -- host-side GPUE equation strings (parser, node array, evaluator)
// and their lowering to the postfix program evaluated per point on the GPU.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace gpue {

struct EqnNodeC {   // EqnNode_gpu (include/ds.h:71-80)
  double val;
  int is_dynamic;
  char var;
  int left, right;
  int op_num;
};

// Postfix program: CONST pushes val; X/Y/Z/T push a coordinate; OP+k pops
// (a, b) and pushes GPUE operator k applied to (a, b) (src/dynamic.cu:568-600).
enum { EQH_CONST = 0, EQH_X = 1, EQH_Y = 2, EQH_Z = 3, EQH_T = 4, EQH_OP = 16 };
constexpr int EQ_MAX_H = 48;
struct EqnProgH {
  int n;
  int depth;   // largest stack depth the program reaches
  int op[EQ_MAX_H];
  double val[EQ_MAX_H];
};

int eqn_parse(const std::string& expr, const std::map<std::string, double>& doubles,
              const std::map<std::string, std::vector<EqnNodeC>>& earlier, std::vector<EqnNodeC>& out,
              bool& time_dep, std::string& err);
double eqn_eval(const EqnNodeC* nodes, int n, double x, double y, double z, double t);
int eqn_to_program(const EqnNodeC* nodes, int n, EqnProgH& prog, std::string& err);

}  // namespace gpue
