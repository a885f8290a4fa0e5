// eqn.cpp -- GPUE equation strings: parser (src/dynamic.cu:227-470),
// pre-order node array (tree_to_array, :472-507) and evaluator (:527-603).
//
// GPUE's grammar is kept exactly, quirks included (they decide which field a
// config file produces):
//   * whitespace is removed by the param-file reader, brackets ( [ ] ) group;
//   * the split point is the leftmost '-', else leftmost '+', else '/', '*',
//     '^' outside brackets, so a-b+c is a-(b+c) and a/b*c is a/(b*c);
//   * literals are digits and '.' only; x y z t are coordinates/time;
//   * cos sin tan sqrt exp take one argument, pow(a,b) = a^(int)b;
//   * function nodes (op >= 6) keep only their left child (tree_to_array),
//     so a second argument of jn/yn is dropped, as in GPUE;
//   * other names are earlier equations (inlined, marking the field time
//     dependent) or Grid doubles.
// k2n is rejected: GPUE's poly_i reads an uninitialised accumulator
// (src/dynamic.cu:113-124), so it has no defined value to reproduce.
#include <cctype>
#include <cmath>
#include <functional>
#include <cstring>
#include <map>
#include <memory>
#include <stack>
#include <stdexcept>
#include <string>
#include <vector>

#include "eqn.h"

#include <algorithm>

namespace gpue {

namespace {

struct Node {
  double val = 0;
  bool is_dynamic = false;
  char var = '0';
  int op_num = 0;
  bool has_op = false;
  std::shared_ptr<Node> left, right;
};

struct ParseCtx {
  const std::map<std::string, double>* doubles;
  const std::map<std::string, std::shared_ptr<Node>>* asts;
};

Node parse(ParseCtx& pc, std::string eqn) {
  if (eqn.empty()) throw std::runtime_error("empty (sub)expression");
  static const char* mops[5] = {"-", "+", "/", "*", "^"};
  auto opnum = [](char ch) {
    switch (ch) {
      case '+': return 1;
      case '-': return 2;
      case '*': return 3;
      case '/': return 4;
      case '^': return 5;
    }
    return 0;
  };
  static const std::map<std::string, int> funcs = {{"cos", 6},  {"sin", 7},  {"tan", 8},  {"sqrt", 9},
                                                   {"exp", 10}, {"pow", 5},  {"jn", 11},  {"yn", 12},
                                                   {"k2n", 13}};
  // bracket pairs in closing order (src/dynamic.cu:282-297)
  std::stack<int> open;
  std::vector<int> ign;
  for (int i = 0; i < (int)eqn.size(); ++i) {
    if (eqn[i] == '(' || eqn[i] == '[') open.push(i);
    if (eqn[i] == ']' || eqn[i] == ')') {
      if (open.empty()) throw std::runtime_error("unbalanced brackets");
      ign.push_back(open.top());
      ign.push_back(i);
      open.pop();
    }
  }
  if (!open.empty()) throw std::runtime_error("unbalanced brackets");
  std::string tmp = eqn;
  if (!ign.empty()) {
    // strip a pair covering the whole expression (:304-316)
    if (ign[ign.size() - 1] == (int)tmp.size() - 1 && ign[ign.size() - 2] == 0) {
      ign.erase(ign.end() - 1);
      ign.erase(ign.end() - 1);
      eqn = eqn.substr(1, eqn.size() - 2);
      for (auto& q : ign) --q;
    }
    tmp = eqn;
    std::vector<int> tp = ign;
    for (size_t i = 0; i < tp.size(); i += 2) {   // (:321-330)
      tmp.erase(tp[i], tp[i + 1] - tp[i] + 1);
      for (size_t j = i + 2; j < tp.size(); ++j)
        if (tp[j] > tp[i]) tp[j] -= tp[i + 1] - tp[i] + 1;
    }
  }
  Node n;
  if (tmp.find_first_not_of("0123456789.") == std::string::npos) {   // (:336-342)
    n.val = std::atof(tmp.c_str());
    return n;
  }
  if (tmp.size() == 1) {   // (:343-365)
    char ch = tmp[0];
    if (ch == 'x' || ch == 'y' || ch == 'z' || ch == 't') {
      n.is_dynamic = true;
      n.var = ch;
      return n;
    }
  }
  int mop = -1;
  for (auto m : mops) {   // (:373-379)
    size_t f = tmp.find(m);
    if (f < tmp.size()) { mop = (int)f; break; }
  }
  if (mop < 0) {
    auto it = funcs.find(tmp);
    if (it != funcs.end()) {   // (:381-409)
      if (it->second == 13) throw std::runtime_error("k2n is not supported (undefined in GPUE)");
      int mp = (int)tmp.size() - 1;
      std::string chk = eqn.substr(mp + 1, eqn.size() - mp - 1);
      n.op_num = it->second;
      n.has_op = true;
      size_t comma = chk.find(",");
      if (comma < chk.size()) {
        n.left = std::make_shared<Node>(parse(pc, chk.substr(1, comma - 1)));
        n.right = std::make_shared<Node>(parse(pc, chk.substr(comma + 1, chk.size() - comma - 2)));
      } else {
        n.left = std::make_shared<Node>(parse(pc, chk));
      }
      return n;
    }
    auto a = pc.asts->find(tmp);   // (:411-425)
    if (a != pc.asts->end()) return *a->second;
    auto d = pc.doubles->find(tmp);
    if (d != pc.doubles->end()) {
      n.val = d->second;
      return n;
    }
    throw std::runtime_error("no value '" + tmp + "' found");
  }
  if (!ign.empty()) {   // map the split point back into eqn (:440-452)
    int count = 0;
    for (int i = 0; i <= mop; ++i) {
      for (size_t j = 0; j < ign.size(); j += 2)
        if (ign[j] == count) count += ign[j + 1] - ign[j] + 1;
      count++;
    }
    mop = count - 1;
  }
  n.op_num = opnum(eqn[mop]);
  n.has_op = true;
  n.left = std::make_shared<Node>(parse(pc, eqn.substr(0, mop)));
  n.right = std::make_shared<Node>(parse(pc, eqn.substr(mop + 1, eqn.size() - mop - 1)));
  return n;
}

// tree_to_array (src/dynamic.cu:472-507)
void to_array(const Node& n, std::vector<EqnNodeC>& out) {
  int me = (int)out.size();
  EqnNodeC e;
  e.val = n.val;
  e.is_dynamic = n.is_dynamic ? 1 : 0;
  e.var = n.var;
  e.op_num = n.op_num;
  e.left = e.right = -1;
  out.push_back(e);
  if (!n.has_op) return;
  out[me].left = (int)out.size();
  to_array(*n.left, out);
  if (n.op_num < 6) {
    if (!n.right) throw std::runtime_error("binary operator without right operand");
    out[me].right = (int)out.size();
    to_array(*n.right, out);
  }
}

double eval_op(int op, double a, double b) {
  switch (op) {
    case 1: return a + b;
    case 2: return a - b;
    case 3: return a * b;
    case 4: return a / b;
    case 5: return std::pow(a, (int)b);
    case 6: return std::cos(a);
    case 7: return std::sin(a);
    case 8: return std::tan(a);
    case 9: return std::sqrt(a);
    case 10: return std::exp(a);
    case 11: { int k = (int)a; return k == 0 ? ::j0(b) : k == 1 ? ::j1(b) : ::jn(k, b); }
    case 12: { int k = (int)a; return k == 0 ? ::y0(b) : k == 1 ? ::y1(b) : ::yn(k, b); }
  }
  return 0.0;
}

double eval_node(const EqnNodeC* e, int i, double x, double y, double z, double t) {
  if (e[i].left < 0 && e[i].right < 0) {
    if (e[i].is_dynamic) {
      switch (e[i].var) {
        case 'x': return x;
        case 'y': return y;
        case 'z': return z;
        case 't': return t;
      }
    }
    return e[i].val;
  }
  double v1 = e[i].left > 0 ? eval_node(e, e[i].left, x, y, z, t) : 0.0;
  double v2 = e[i].right > 0 ? eval_node(e, e[i].right, x, y, z, t) : 0.0;
  return eval_op(e[i].op_num, v1, v2);
}

void to_postfix(const EqnNodeC* e, int i, EqnProgH& p) {
  auto emit = [&](int op, double v) {
    if (p.n >= EQ_MAX_H) throw std::runtime_error("equation too long for the device program");
    p.op[p.n] = op;
    p.val[p.n] = v;
    p.n++;
  };
  if (e[i].left < 0 && e[i].right < 0) {
    if (e[i].is_dynamic) {
      switch (e[i].var) {
        case 'x': emit(EQH_X, 0); return;
        case 'y': emit(EQH_Y, 0); return;
        case 'z': emit(EQH_Z, 0); return;
        case 't': emit(EQH_T, 0); return;
      }
    }
    emit(EQH_CONST, e[i].val);
    return;
  }
  if (e[i].left > 0) to_postfix(e, e[i].left, p);
  else emit(EQH_CONST, 0.0);
  if (e[i].right > 0) to_postfix(e, e[i].right, p);
  else emit(EQH_CONST, 0.0);
  emit(EQH_OP + e[i].op_num, 0.0);
}

}  // namespace

int eqn_parse(const std::string& expr, const std::map<std::string, double>& doubles,
              const std::map<std::string, std::vector<EqnNodeC>>& earlier, std::vector<EqnNodeC>& out,
              bool& time_dep, std::string& err) {
  try {
    // rebuild earlier equations as trees so that references inline them
    std::map<std::string, std::shared_ptr<Node>> asts;
    for (auto& kv : earlier) {
      std::function<std::shared_ptr<Node>(int)> build = [&](int i) {
        auto n = std::make_shared<Node>();
        const EqnNodeC& e = kv.second[i];
        n->val = e.val;
        n->is_dynamic = e.is_dynamic != 0;
        n->var = e.var;
        n->op_num = e.op_num;
        n->has_op = e.left >= 0;
        if (e.left >= 0) n->left = build(e.left);
        if (e.right >= 0) n->right = build(e.right);
        return n;
      };
      if (!kv.second.empty()) asts[kv.first] = build(0);
    }
    std::string s;
    for (char ch : expr)
      if (!std::isspace((unsigned char)ch)) s.push_back(ch);
    ParseCtx pc{&doubles, &asts};
    Node root = parse(pc, s);
    out.clear();
    to_array(root, out);
    time_dep = false;   // a field is time dependent iff a 't' leaf survives inlining
    for (auto& e : out)
      if (e.is_dynamic && e.var == 't') time_dep = true;
    return 0;
  } catch (const std::exception& ex) {
    err = ex.what();
    return 1;
  }
}

double eqn_eval(const EqnNodeC* nodes, int n, double x, double y, double z, double t) {
  if (n <= 0) return 0.0;
  return eval_node(nodes, 0, x, y, z, t);
}

int eqn_to_program(const EqnNodeC* nodes, int n, EqnProgH& prog, std::string& err) {
  try {
    prog.n = 0;
    if (n > 0) to_postfix(nodes, 0, prog);
    int sp = 0;
    prog.depth = 0;
    for (int i = 0; i < prog.n; ++i) {
      sp += prog.op[i] < EQH_OP ? 1 : -1;
      prog.depth = std::max(prog.depth, sp);
    }
    return 0;
  } catch (const std::exception& ex) {
    err = ex.what();
    return 1;
  }
}

}  // namespace gpue
