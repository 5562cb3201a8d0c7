// expr.hip — compiler of the reference's coefficient expressions into the
// register bytecode of expr.hpp, and mhx_expr_eval (host evaluation of a
// compiled function, the role of FunctionManager::evaluate at "point").
//
// Reference semantics (parsing: interpreter.cpp:63-352, operator form
// :359-452; leaf resolution: functionManager_create.hpp:95-500; evaluation:
// functionManager_evaluate.hpp:56-1353):
//   - a string is split at its top-level '+'/'-' if it has any, else at
//     '*', '/', '<', '>', '<=', '>=', else at '^'; the pieces of a level are
//     folded left to right (a-b+c = (a-b)+c, a^b^c = (a^b)^c), so '<' binds
//     tighter than '+' and '^' tighter than '*';
//   - while splitting at '+'/'-' every space and '=' of the string is dropped;
//     at the product level spaces are dropped; a '+'/'-' splits only after a
//     non-empty piece (so "2e-3" inside a sum splits at its '-');
//   - a leading '-' reads as "0.0-...";
//   - leaves, in this order: solution fields ("T", "grad(T)[x|y|z]", "T_t"),
//     scalar fields ("x", "y", "z", and "u", "v", "w", zero without mesh
//     motion), element data ("k0", "src" where given), other functions by
//     name, numbers, "t", "pi";
//   - name(arg) and name(a,b) for the operators sin cos exp log tan abs max
//     min mean emax emin emean sqrt sinh cosh; comparisons give 1 or 0;
//     sqrt of a non-positive value is 0; mean(a,b) = 0.5 a + 0.5 b;
//   - a power whose exponent is a scalar (no field in it) and whose base
//     branch varies over the points is applied as repeated products when the
//     exponent is within 1e-12 of 2, 3 or 4, else as pow;
//   - branches of numbers and pi only are evaluated once, on the host.
// Forms the reference accepts but evaluates to undefined data (an unknown
// operator name, two-argument operators with parentheses inside their
// arguments, "h", element reductions of fields) are refused here with a
// message.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "expr.hpp"
#include "mhx_internal.hpp"

namespace mhx {

namespace {

// value classes of a branch (functionManager_create.hpp:587-619): constant
// (numbers, pi), scalar (time), a per-point view, a view with derivatives
enum XType { XT_CONST = 0, XT_SCALAR = 1, XT_VIEW = 2, XT_AD = 3 };

enum XOpName {
  ON_ASSIGN = 0, ON_PLUS, ON_MINUS, ON_TIMES, ON_DIVIDE, ON_POWER, ON_LT, ON_LTE, ON_GT, ON_GTE,
  ON_SIN, ON_COS, ON_EXP, ON_LOG, ON_TAN, ON_ABS, ON_MAX, ON_MIN, ON_MEAN, ON_EMAX, ON_EMIN, ON_EMEAN,
  ON_SQRT, ON_SINH, ON_COSH
};

// the operator names, in the order the reference tries them
// (functionManager_construct.hpp:34)
struct OpName {
  const char* name;
  XOpName op;
};
const OpName kOps[] = {{"sin", ON_SIN},   {"cos", ON_COS},     {"exp", ON_EXP},   {"log", ON_LOG},
                       {"tan", ON_TAN},   {"abs", ON_ABS},     {"max", ON_MAX},   {"min", ON_MIN},
                       {"mean", ON_MEAN}, {"emax", ON_EMAX},   {"emin", ON_EMIN}, {"emean", ON_EMEAN},
                       {"sqrt", ON_SQRT}, {"sinh", ON_SINH},   {"cosh", ON_COSH}};

bool is_binary_fn(XOpName o) { return o == ON_MAX || o == ON_MIN || o == ON_MEAN; }
bool is_unary_fn(XOpName o) { return o >= ON_SIN && !is_binary_fn(o); }

struct XNode {
  bool leaf = false;
  double c = 0.0;      // constant leaf or folded value
  int var = -1;        // variable leaf: ExprVar
  XType type = XT_CONST;
  std::vector<int> dep;
  std::vector<XOpName> op;
};

// a number in the reference's sense (interpreter.cpp:18-56): digits, at most
// one '.', at most one 'e'/'E', '+'/'-' only first or after 'e'/'E'
bool is_number(const std::string& s) {
  int dots = 0, es = 0;
  bool ok = true;
  for (size_t k = 0; k < s.size(); ++k) {
    const char ch = s[k];
    if (ch >= '0' && ch <= '9') continue;
    if (ch == '.') {
      if (++dots > 1) ok = false;
    } else if (ch == 'e' || ch == 'E') {
      if (++es > 1) ok = false;
    } else if (ch == '+' || ch == '-') {
      if (k > 0 && s[k - 1] != 'e' && s[k - 1] != 'E') ok = false;
    } else {
      ok = false;
    }
  }
  return ok;
}

double host_op(XOpName o, double a, double b) {
  switch (o) {
    case ON_ASSIGN: return b;
    case ON_PLUS: return a + b;
    case ON_MINUS: return a + (-b);
    case ON_TIMES: return a * b;
    case ON_DIVIDE: return a / b;
    case ON_POWER: return std::pow(a, b);
    case ON_LT: return a < b ? 1.0 : 0.0;
    case ON_LTE: return a <= b ? 1.0 : 0.0;
    case ON_GT: return a > b ? 1.0 : 0.0;
    case ON_GTE: return a >= b ? 1.0 : 0.0;
    case ON_SIN: return std::sin(b);
    case ON_COS: return std::cos(b);
    case ON_EXP: return std::exp(b);
    case ON_LOG: return std::log(b);
    case ON_TAN: return std::tan(b);
    case ON_ABS: return b < 0.0 ? -b : b;
    case ON_MAX: return a < b ? b : a;
    case ON_MIN: return b < a ? b : a;
    case ON_MEAN: return 0.5 * a + 0.5 * b;
    case ON_EMAX: case ON_EMIN: case ON_EMEAN: return b;  // of a scalar: itself
    case ON_SQRT: return b <= 0.0 ? 0.0 : std::sqrt(b);
    case ON_SINH: return std::sinh(b);
    case ON_COSH: return std::cosh(b);
  }
  return 0.0;
}

class Compiler {
 public:
  Compiler(const std::vector<std::string>& names, const std::vector<std::string>& exprs, bool have_k0,
           bool have_src)
      : names_(names), exprs_(exprs), have_k0_(have_k0), have_src_(have_src), root_(names.size(), -2) {}

  // root node of function f (parsed once; -2: not yet, -3: in progress)
  int function(size_t f) {
    if (root_[f] == -3) fail("cyclic reference through function '" + names_[f] + "'");
    if (root_[f] >= 0) return root_[f];
    root_[f] = -3;
    const int r = parse(exprs_[f], 0);
    root_[f] = r;
    return r;
  }

  // emit function f's program; returns its length
  int emit(size_t f, std::vector<ExprIns>& code, std::vector<double>& consts) {
    code_ = &code;
    consts_ = &consts;
    const size_t start = code.size();
    gen(function(f), 0);
    return (int)(code.size() - start);
  }

  const XNode& node(int i) const { return nodes_[i]; }

 private:
  [[noreturn]] void fail(const std::string& m) { throw Error{MHX_ERR_ARG, "expression: " + m}; }

  int add(XNode n) {
    nodes_.push_back(std::move(n));
    return (int)nodes_.size() - 1;
  }
  int constant(double v) {
    XNode n;
    n.leaf = true;
    n.c = v;
    n.type = XT_CONST;
    return add(n);
  }
  int variable(ExprVar v, XType t) {
    XNode n;
    n.leaf = true;
    n.var = v;
    n.type = t;
    return add(n);
  }

  // a fold node from its pieces; branches of constants only are evaluated
  // here (SToS semantics, functionManager_evaluate.hpp:1243-1353)
  int fold(std::vector<int> dep, std::vector<XOpName> op) {
    XNode n;
    bool all_const = true;
    int t = XT_CONST;
    for (int d : dep) {
      all_const = all_const && nodes_[d].type == XT_CONST;
      t = std::max(t, (int)nodes_[d].type);
    }
    if (all_const) {
      double v = 0.0;
      for (size_t k = 0; k < dep.size(); ++k) v = host_op(op[k], v, nodes_[dep[k]].c);
      return constant(v);
    }
    n.type = (XType)t;
    n.dep = std::move(dep);
    n.op = std::move(op);
    return add(n);
  }

  int parse(const std::string& s0, int depth) {
    if (depth > 200) fail("nesting too deep in '" + s0 + "'");
    // leaves (functionManager_create.hpp:105-459)
    if (s0 == "T") return variable(XV_T, XT_AD);
    if (s0 == "grad(T)[x]") return variable(XV_GX, XT_AD);
    if (s0 == "grad(T)[y]") return variable(XV_GY, XT_AD);
    if (s0 == "grad(T)[z]") return variable(XV_GZ, XT_AD);
    if (s0 == "T_t") return variable(XV_TT, XT_AD);
    if (s0 == "x") return variable(XV_X, XT_VIEW);
    if (s0 == "y") return variable(XV_Y, XT_VIEW);
    if (s0 == "z") return variable(XV_Z, XT_VIEW);
    if (s0 == "u" || s0 == "v" || s0 == "w") return variable(XV_ZERO, XT_VIEW);
    if (s0 == "k0" && have_k0_) return variable(XV_K0, XT_VIEW);
    if (s0 == "src" && have_src_) return variable(XV_SRCE, XT_VIEW);
    for (size_t f = 0; f < names_.size(); ++f)
      if (s0 == names_[f]) return function(f);
    if (!s0.empty() && is_number(s0)) {
      try {
        return constant(std::stod(s0));
      } catch (...) {
        fail("'" + s0 + "' is not a number");
      }
    }
    if (s0 == "t") return variable(XV_TIME, XT_SCALAR);
    if (s0 == "pi") return constant(3.141592653589793238463);
    if (s0 == "h") fail("'h' has no value at volume points");
    // operator form name(arg) / name(a,b) (interpreter.cpp:359-452): the
    // name, '(' right after it, ')' last and nowhere before
    for (const OpName& on : kOps) {
      const size_t L = std::strlen(on.name);
      if (s0.size() < L + 2 || s0.compare(0, L, on.name) != 0) continue;
      if (s0[L] != '(' || s0.back() != ')') continue;
      if (s0.find(')', L + 1) != s0.size() - 1) continue;
      const std::string arg = s0.substr(L + 1, s0.size() - L - 2);
      size_t comma = std::string::npos;
      for (size_t i = 0; i + 1 < arg.size(); ++i)
        if (arg[i] == ',') comma = i;
      if (comma != std::string::npos) {
        if (!is_binary_fn(on.op)) fail("operator '" + std::string(on.name) + "' takes one argument");
        const int a = parse(arg.substr(0, comma), depth + 1), b = parse(arg.substr(comma + 1), depth + 1);
        return fold({a, b}, {ON_ASSIGN, on.op});
      }
      if (is_binary_fn(on.op)) fail("operator '" + std::string(on.name) + "' takes two arguments");
      return unary(on.op, parse(arg, depth + 1));
    }
    return split(s0, depth);
  }

  int unary(XOpName o, int a) {
    if ((o == ON_EMAX || o == ON_EMIN || o == ON_EMEAN) && nodes_[a].type >= XT_VIEW)
      fail("element reductions (emax, emin, emean) of point fields are not built");
    return fold({a}, {o});
  }

  // interpreter.cpp:63-352
  int split(std::string s, int depth) {
    if (!s.empty() && s[0] == '-') s = "0.0" + s;
    if (s.size() <= 1) fail("could not decompose or find '" + s + "'");
    int npm = 0, nmdp = 0, npow = 0, par = 0;
    for (char ch : s) {
      if (ch == '(') ++par;
      else if (ch == ')') --par;
      else if (par == 0) {
        npm += ch == '+' || ch == '-';
        nmdp += ch == '*' || ch == '/' || ch == '<' || ch == '>';
        npow += ch == '^';
      }
    }
    if (par != 0) fail(std::string(par > 0 ? "unclosed" : "extra") + " parenthesis in '" + s + "'");
    std::vector<std::string> piece;
    std::vector<XOpName> op;
    std::string cur;
    XOpName curop = ON_ASSIGN;
    par = 0;
    if (npm > 0) {
      for (char ch : s) {
        if (ch == ' ' || ch == '=') continue;
        if (ch == '(') ++par;
        if (ch == ')') --par;
        if (par == 0 && (ch == '+' || ch == '-') && !cur.empty()) {
          piece.push_back(cur), op.push_back(curop);
          cur.clear();
          curop = ch == '+' ? ON_PLUS : ON_MINUS;
          continue;
        }
        cur += ch;
      }
      if (!cur.empty()) piece.push_back(cur), op.push_back(curop);
    } else if (nmdp > 0) {
      for (size_t i = 0; i < s.size(); ++i) {
        const char ch = s[i];
        if (ch == ' ') continue;
        if (ch == '(') ++par;
        if (ch == ')') --par;
        if (par == 0 && (ch == '*' || ch == '/' || ch == '<' || ch == '>')) {
          piece.push_back(cur), op.push_back(curop);
          cur.clear();
          const bool eq = i + 1 < s.size() && s[i + 1] == '=';
          if (ch == '*') curop = ON_TIMES;
          else if (ch == '/') curop = ON_DIVIDE;
          else if (ch == '<') curop = eq ? ON_LTE : ON_LT;
          else curop = eq ? ON_GTE : ON_GT;
          if ((ch == '<' || ch == '>') && eq) ++i;
          continue;
        }
        cur += ch;
      }
      piece.push_back(cur), op.push_back(curop);
    } else if (npow > 0) {
      for (char ch : s) {
        if (ch == '(') ++par;
        if (ch == ')') --par;
        if (par == 0 && ch == '^') {
          piece.push_back(cur), op.push_back(curop);
          cur.clear();
          curop = ON_POWER;
          continue;
        }
        cur += ch;
      }
      piece.push_back(cur), op.push_back(curop);
    } else if (s[0] == '(' && s.back() == ')') {
      return parse(s.substr(1, s.size() - 2), depth + 1);
    } else {
      // name(argument) with parentheses inside the argument
      const size_t p = s.find('(', 1);
      if (p == std::string::npos || p >= s.size() - 1 || s.back() != ')')
        fail("could not decompose or find '" + s + "'");
      const std::string name = s.substr(0, p), arg = s.substr(p + 1, s.size() - p - 2);
      for (const OpName& on : kOps)
        if (name == on.name) {
          if (is_binary_fn(on.op)) fail("operator '" + name + "' with parentheses inside its arguments");
          return unary(on.op, parse(arg, depth + 1));
        }
      fail("unknown operator '" + name + "' in '" + s + "'");
    }
    std::vector<int> dep;
    for (size_t k = 0; k < piece.size(); ++k) {
      if (piece[k].empty()) fail("empty operand in '" + s + "'");
      dep.push_back(parse(piece[k], depth + 1));
    }
    return fold(dep, op);
  }

  void emit1(ExprOp o, int dst, int a, int b) {
    if (code_->size() >= (size_t)kExprMaxCode) fail("programs longer than the evaluator allows");
    code_->push_back(ExprIns{(uint8_t)o, (uint8_t)dst, (uint8_t)a, (uint8_t)b});
  }

  // node n into register r (registers r.. are free)
  void gen(int n, int r) {
    if (r >= kExprRegs) fail("expression nested deeper than the evaluator's registers");
    const XNode& x = nodes_[n];
    if (x.type == XT_CONST) {
      if (consts_->size() >= (size_t)kExprMaxConst) fail("too many constants");
      const int ci = (int)consts_->size();
      consts_->push_back(x.c);
      emit1(XOP_CONST, r, ci & 0xff, ci >> 8);
      return;
    }
    if (x.leaf) {
      emit1(XOP_VAR, r, x.var, 0);
      return;
    }
    const bool view = x.type >= XT_VIEW;
    for (size_t k = 0; k < x.dep.size(); ++k) {
      const XOpName o = x.op[k];
      const int d = x.dep[k];
      if (k == 0 && o == ON_ASSIGN) {
        gen(d, r);
        continue;
      }
      if (k == 0 || is_unary_fn(o)) {
        // a unary operator (or an operator first in a fold, which the
        // reference applies to an unset value): r = f(dep)
        if (!is_unary_fn(o)) fail("operator without a left operand");
        if (o == ON_EMAX || o == ON_EMIN || o == ON_EMEAN) {
          gen(d, r);  // of a scalar: itself
          continue;
        }
        gen(d, r);
        emit1(unary_code(o), r, r, r);
        continue;
      }
      gen(d, r + 1);
      ExprOp code;
      switch (o) {
        case ON_PLUS: code = XOP_ADD; break;
        case ON_MINUS: code = XOP_SUB; break;
        case ON_TIMES: code = XOP_MUL; break;
        case ON_DIVIDE: code = XOP_DIV; break;
        case ON_POWER: code = (view && nodes_[d].type <= XT_SCALAR) ? XOP_POWS : XOP_POW; break;
        case ON_LT: code = XOP_LT; break;
        case ON_LTE: code = XOP_LTE; break;
        case ON_GT: code = XOP_GT; break;
        case ON_GTE: code = XOP_GTE; break;
        case ON_MAX: code = XOP_MAX; break;
        case ON_MIN: code = XOP_MIN; break;
        case ON_MEAN: code = XOP_MEAN; break;
        default: fail("misplaced operator");
      }
      emit1(code, r, r, r + 1);
    }
  }

  static ExprOp unary_code(XOpName o) {
    switch (o) {
      case ON_SIN: return XOP_SIN;
      case ON_COS: return XOP_COS;
      case ON_TAN: return XOP_TAN;
      case ON_EXP: return XOP_EXP;
      case ON_LOG: return XOP_LOG;
      case ON_ABS: return XOP_ABS;
      case ON_SQRT: return XOP_SQRT;
      case ON_SINH: return XOP_SINH;
      default: return XOP_COSH;
    }
  }

  const std::vector<std::string>& names_;
  const std::vector<std::string>& exprs_;
  bool have_k0_, have_src_;
  std::vector<int> root_;
  std::vector<XNode> nodes_;
  std::vector<ExprIns>* code_ = nullptr;
  std::vector<double>* consts_ = nullptr;
};

// coefficient names of the thermal module (thermal.cpp:55-61) and their
// defaults
struct FnName {
  const char* name;
  double dflt;
};
const FnName kThermalFns[XF_COUNT] = {{"thermal diffusion", 1.0}, {"thermal source", 0.0}, {"specific heat", 1.0},
                                      {"density", 1.0},           {"advection x", 0.0},    {"advection y", 0.0},
                                      {"advection z", 0.0}};

}  // namespace

void expr_compile(const mhx_problem_desc& d, ExprHost& out) {
  out = ExprHost{};
  for (int f = 0; f < XF_COUNT; ++f) out.dflt[f] = kThermalFns[f].dflt;
  if (d.n_functions <= 0) return;
  MHX_REQUIRE(d.function_names && d.function_exprs, MHX_ERR_ARG, "n_functions > 0 needs function_names and exprs");
  std::vector<std::string> names, exprs;
  for (int i = 0; i < d.n_functions; ++i) {
    MHX_REQUIRE(d.function_names[i] && d.function_exprs[i], MHX_ERR_ARG, "null function name or expression");
    names.emplace_back(d.function_names[i]);
    exprs.emplace_back(d.function_exprs[i]);
  }
  Compiler comp(names, exprs, d.k0 != nullptr, d.src_elem != nullptr);
  for (int f = 0; f < XF_COUNT; ++f) {
    for (size_t i = 0; i < names.size(); ++i)
      if (names[i] == kThermalFns[f].name) {
        out.off[f] = (int)out.code.size();
        out.len[f] = comp.emit(i, out.code, out.consts);
        out.active = 1;
      }
  }
  // every function compiles, also those no coefficient reads (errors surface
  // at creation, as the reference decomposes every registered function)
  std::vector<ExprIns> scratch_code;
  std::vector<double> scratch_consts;
  for (size_t i = 0; i < names.size(); ++i) comp.emit(i, scratch_code, scratch_consts);
}

}  // namespace mhx

using namespace mhx;

static thread_local std::string g_expr_err;

extern "C" int mhx_expr_eval(const mhx_problem_desc* desc, const char* fname, int64_t n, const double* pts,
                             double* out) {
  g_expr_err.clear();
  try {
    if (!desc || !fname || n < 0 || (n > 0 && (!pts || !out))) return MHX_ERR_ARG;
    std::vector<std::string> names, exprs;
    int which = -1;
    for (int i = 0; i < desc->n_functions; ++i) {
      if (!desc->function_names || !desc->function_exprs || !desc->function_names[i] || !desc->function_exprs[i])
        return MHX_ERR_ARG;
      names.emplace_back(desc->function_names[i]);
      exprs.emplace_back(desc->function_exprs[i]);
      if (names.back() == fname) which = i;
    }
    if (which < 0) {
      g_expr_err = std::string("no function named '") + fname + "'";
      return MHX_ERR_ARG;
    }
    Compiler comp(names, exprs, true, true);
    std::vector<ExprIns> code;
    std::vector<double> consts;
    const int len = comp.emit((size_t)which, code, consts);
    // point record: T, Tt, grad T (3), x (3), t, k0, src -> value and
    // d/dT, d/dgrad_x, d/dgrad_y, d/dgrad_z, d/dTt
    for (int64_t i = 0; i < n; ++i) {
      const double* q = pts + 11 * i;
      ExprPoint<5> p;
      p.T = xq_const<5>(q[0]);
      p.T.d[0] = 1.0;
      p.Tt = xq_const<5>(q[1]);
      p.Tt.d[4] = 1.0;
      for (int c = 0; c < 3; ++c) {
        p.g[c] = xq_const<5>(q[2 + c]);
        p.g[c].d[1 + c] = 1.0;
        p.x[c] = q[5 + c];
      }
      p.time = q[8];
      p.k0 = q[9];
      p.src = q[10];
      const QDual<5> r = expr_run<5>(code.data(), consts.data(), 0, len, p);
      out[6 * i] = r.v;
      for (int k = 0; k < 5; ++k) out[6 * i + 1 + k] = r.d[k];
    }
    return MHX_OK;
  } catch (const Error& e) {
    g_expr_err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    g_expr_err = e.what();
    return MHX_ERR_ARG;
  }
}

extern "C" const char* mhx_expr_last_error(void) { return g_expr_err.c_str(); }
