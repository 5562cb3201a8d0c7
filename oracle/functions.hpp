// functions.hpp — CPU restatement of the reference's FunctionManager and
// Interpreter (TEST INFRASTRUCTURE: part of the oracle, included by
// oracle.cpp only).
//
// Follows, step by step:
//   Interpreter::isScalar                  src/tools/interpreter.cpp:18-56
//   Interpreter::split                     src/tools/interpreter.cpp:63-352
//   Interpreter::isOperator                src/tools/interpreter.cpp:359-452
//   FunctionManager::addFunction           functionManager_create.hpp:14-39
//   FunctionManager::decomposeFunctions    functionManager_create.hpp:61-550
//     (leaf order: solution fields, scalar fields, parameters, functions,
//      numbers, known variables, operators, split)
//   FunctionManager::checkDepDataType      functionManager_create.hpp:587-619
//   FunctionManager::evaluate              functionManager_evaluate.hpp:56-225
//   evaluateOpVToV / SToV / SToS           functionManager_evaluate.hpp:237-569,
//                                          :891-1242, :1243-1353
// Branches are kept in the reference's flat list per function ("tree"), each
// with its expression string, dependencies and their operator names; a
// branch is evaluated over the points of one element at a time: a view is a
// vector over the element's points, a scalar a single value.
#pragma once

#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

// Sacado Fad elementary functions (value + derivative slots)
template <int N> inline SFad<N> fchain(const SFad<N>& a, double fv, double fp) {
  SFad<N> r; r.v = fv; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * fp; return r;
}
template <int N> inline SFad<N> sin(const SFad<N>& a) { return fchain(a, std::sin(a.v), std::cos(a.v)); }
template <int N> inline SFad<N> cos(const SFad<N>& a) { return fchain(a, std::cos(a.v), -std::sin(a.v)); }
template <int N> inline SFad<N> tan(const SFad<N>& a) { double t = std::tan(a.v); return fchain(a, t, 1.0 + t * t); }
template <int N> inline SFad<N> exp(const SFad<N>& a) { double e = std::exp(a.v); return fchain(a, e, e); }
template <int N> inline SFad<N> sinh(const SFad<N>& a) { return fchain(a, std::sinh(a.v), std::cosh(a.v)); }
template <int N> inline SFad<N> cosh(const SFad<N>& a) { return fchain(a, std::cosh(a.v), std::sinh(a.v)); }
template <int N> inline SFad<N> log(const SFad<N>& a) {
  SFad<N> r; r.v = std::log(a.v); for (int i = 0; i < N; ++i) r.d[i] = a.d[i] / a.v; return r;
}
// pow: d = pow(a, b) (b da / a + log(a) db), no derivative where a = 0; terms
// with a zero derivative slot are left out
template <int N> inline SFad<N> pow(const SFad<N>& a, const SFad<N>& b) {
  SFad<N> r; r.v = std::pow(a.v, b.v);
  if (a.v == 0.0) { for (int i = 0; i < N; ++i) r.d[i] = 0.0; return r; }
  const double la = std::log(a.v);
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
    if (a.d[i] != 0.0) s = b.v * a.d[i] / a.v;
    if (b.d[i] != 0.0) s = s + la * b.d[i];
    r.d[i] = r.v * s;
  }
  return r;
}
inline double sin(double a) { return std::sin(a); }
inline double cos(double a) { return std::cos(a); }
inline double tan(double a) { return std::tan(a); }
inline double exp(double a) { return std::exp(a); }
inline double log(double a) { return std::log(a); }
inline double sinh(double a) { return std::sinh(a); }
inline double cosh(double a) { return std::cosh(a); }
inline double pow(double a, double b) { return std::pow(a, b); }

struct FBranch {
  std::string expr;
  bool leaf = false, decomposed = false;
  // leaf kinds
  enum Kind { NONE, NUMBER, PI, TIME, FIELD, SCALARF, DATA, FUNC } kind = NONE;
  int index = -1;        // FIELD: 0 T, 1-3 grad(T)[x|y|z], 4 T_t; SCALARF: 0-2 x y z, 3 zero; DATA: 0 k0, 1 src; FUNC: tree
  double cval = 0.0;
  bool is_const = true, is_view = false, is_ad = false;
  std::vector<int> deps;
  std::vector<std::string> ops;
};

struct FTree {
  std::string name;
  std::vector<FBranch> b;
};

[[noreturn]] inline void ffail(const std::string& m) { throw std::runtime_error("function manager: " + m); }

// Interpreter::isScalar
inline bool f_is_scalar(const std::string& s) {
  bool isnum = true;
  int dots = 0, es = 0;
  for (size_t k = 0; k < s.length(); ++k) {
    if (isdigit((unsigned char)s[k])) continue;
    if (s[k] == '.') {
      if (dots == 0) dots = 1; else isnum = false;
    } else if (s[k] == 'e' || s[k] == 'E') {
      if (es == 0) es = 1; else isnum = false;
    } else if (s[k] == '+' || s[k] == '-') {
      if (k > 0 && !(s[k - 1] == 'e' || s[k - 1] == 'E')) isnum = false;
    } else {
      isnum = false;
    }
  }
  return isnum;
}

// Interpreter::split: appends the pieces of branch `idx` as new branches
inline void f_split(std::vector<FBranch>& br, size_t idx) {
  std::string s = br[idx].expr;
  auto push = [&](const std::string& piece, const std::string& op) {
    FBranch nb;
    nb.expr = piece;
    br.push_back(nb);
    br[idx].deps.push_back((int)br.size() - 1);
    br[idx].ops.push_back(op);
  };
  if (!s.empty() && s[0] == '-') s = "0.0" + s;
  if (s.empty()) return;
  if (s.length() == 1) {
    push(std::string(1, s[0]), "");
    return;
  }
  size_t num_pm = 0, num_mdp = 0, num_pow = 0;
  int paren = 0;
  for (size_t i = 0; i < s.length(); ++i) {
    if (s[i] == '(') paren += 1;
    else if (s[i] == ')') paren -= 1;
    else if (paren == 0) {
      if (s[i] == '+' || s[i] == '-') num_pm += 1;
      if (s[i] == '*' || s[i] == '/' || s[i] == '<' || s[i] == '>') num_mdp += 1;
      if (s[i] == '^') num_pow += 1;
    }
  }
  paren = 0;
  std::string cb, cop;
  if (num_pm > 0) {
    for (size_t i = 0; i < s.length(); ++i) {
      if (s[i] == ' ' || s[i] == '=') {
      } else if (s[i] == '(') {
        paren += 1; cb += s[i];
      } else if (s[i] == ')') {
        paren -= 1; cb += s[i];
      } else if (paren == 0 && s[i] == '+' && cb.length() > 0) {
        push(cb, cop); cb = ""; cop = "plus";
      } else if (paren == 0 && s[i] == '-' && cb.length() > 0) {
        push(cb, cop); cb = ""; cop = "minus";
      } else {
        cb += s[i];
      }
      if (i == s.length() - 1 && cb.length() > 0) push(cb, cop);
    }
  } else if (num_mdp > 0) {
    for (size_t i = 0; i < s.length(); ++i) {
      if (s[i] == ' ') {
      } else if (s[i] == '(') {
        paren += 1; cb += s[i];
      } else if (s[i] == ')') {
        paren -= 1; cb += s[i];
      } else if (paren == 0 && s[i] == '*') {
        push(cb, cop); cb = ""; cop = "times";
      } else if (paren == 0 && s[i] == '/') {
        push(cb, cop); cb = ""; cop = "divide";
      } else if (paren == 0 && s[i] == '<') {
        push(cb, cop); cb = "";
        if (i + 1 < s.length() && s[i + 1] == '=') { cop = "lte"; ++i; } else cop = "lt";
      } else if (paren == 0 && s[i] == '>') {
        push(cb, cop); cb = "";
        if (i + 1 < s.length() && s[i + 1] == '=') { cop = "gte"; ++i; } else cop = "gt";
      } else {
        cb += s[i];
      }
      if (i == s.length() - 1) push(cb, cop);
    }
  } else if (num_pow > 0) {
    for (size_t i = 0; i < s.length(); ++i) {
      if (s[i] == '(') {
        paren += 1; cb += s[i];
      } else if (s[i] == ')') {
        paren -= 1; cb += s[i];
      } else if (paren == 0 && s[i] == '^') {
        push(cb, cop); cb = ""; cop = "power";
      } else {
        cb += s[i];
      }
      if (i == s.length() - 1) push(cb, cop);
    }
  } else {
    if (s[0] == '(' && s[s.length() - 1] == ')') {
      push(s.substr(1, s.length() - 2), cop);
    } else {
      size_t pindex = 0;
      bool found = false;
      for (size_t k = 1; k + 1 < s.length(); ++k)
        if (s[k] == '(' && !found) { found = true; pindex = k; }
      if (found && s[s.length() - 1] == ')') push(s.substr(pindex + 1, s.length() - pindex - 2), s.substr(0, pindex));
    }
  }
  if (paren > 0) ffail("unclosed parenthesis in: " + s);
  if (paren < 0) ffail("extra parenthesis in: " + s);
}

inline const std::vector<std::string>& f_known_ops() {
  static const std::vector<std::string> ops = {"sin", "cos", "exp", "log", "tan", "abs", "max", "min",
                                               "mean", "emax", "emin", "emean", "sqrt", "sinh", "cosh"};
  return ops;
}

// Interpreter::isOperator
inline bool f_is_operator(std::vector<FBranch>& br, size_t idx) {
  const std::string s = br[idx].expr;
  const auto& ops = f_known_ops();
  std::string oper, arg;
  bool found = false;
  for (size_t k = 0; !found && k < ops.size(); ++k) {
    const size_t L = ops[k].length();
    if (s.length() < L + 1) continue;
    bool eq = s.compare(0, L, ops[k]) == 0;
    if (eq) {
      if (s[L] != '(' || s[s.length() - 1] != ')') eq = false;
      if (s[L] == '(')
        for (size_t j = L + 1; j + 1 < s.length(); ++j)
          if (s[j] == ')') eq = false;
    }
    if (eq) {
      found = true;
      oper = ops[k];
      arg = s.substr(L + 1, s.length() - L - 2);
    }
  }
  if (!found) return false;
  size_t comma = std::string::npos;
  for (size_t i = 0; i + 1 < arg.length(); ++i)
    if (arg[i] == ',') comma = i;
  auto push = [&](const std::string& piece, const std::string& op) {
    FBranch nb;
    nb.expr = piece;
    br.push_back(nb);
    br[idx].deps.push_back((int)br.size() - 1);
    br[idx].ops.push_back(op);
  };
  if (comma != std::string::npos) {
    push(arg.substr(0, comma), "");
    push(arg.substr(comma + 1), oper);
  } else {
    push(arg, oper);
  }
  br[idx].decomposed = true;
  return true;
}

// values of one element's points: views per point, scalars
template <class E>
struct FElem {
  int nq = 0;
  const E* T = nullptr;       // [nq]
  const E* dT = nullptr;      // [nq * dim]
  const E* Tt = nullptr;      // [nq] (nullptr: steady, zero)
  const double* x = nullptr;  // [nq * 3]
  int dim = 2;
  double time = 0.0, k0 = 1.0, src = 0.0;
};

// a branch value: a view over the element's points or a scalar
template <class E>
struct FVal {
  bool view = false;
  std::vector<E> v;   // view
  E s = E(0.0);       // scalar
};

struct FManager {
  std::vector<FTree> trees;
  bool have_k0 = false, have_src = false;

  void add(const std::string& name, const std::string& expr) {
    FTree t;
    t.name = name;
    FBranch b;
    b.expr = expr;
    t.b.push_back(b);
    trees.push_back(t);
  }
  int find(const std::string& name) const {
    for (size_t i = 0; i < trees.size(); ++i)
      if (trees[i].name == name) return (int)i;
    return -1;
  }

  // decomposeFunctions
  void decompose() {
    for (size_t ti = 0; ti < trees.size(); ++ti) {
      bool done = false;
      int iter = 0;
      while (!done && iter < 100) {
        ++iter;
        auto& br = trees[ti].b;
        for (size_t k = 0; k < br.size(); ++k) {
          if (br[k].leaf || br[k].decomposed) continue;
          const std::string e = br[k].expr;
          FBranch& b = br[k];
          auto leaf = [&](FBranch::Kind kind, int index) {
            br[k].leaf = true;
            br[k].decomposed = true;
            br[k].kind = kind;
            br[k].index = index;
          };
          if (e == "T") { leaf(FBranch::FIELD, 0); continue; }
          if (e == "grad(T)[x]") { leaf(FBranch::FIELD, 1); continue; }
          if (e == "grad(T)[y]") { leaf(FBranch::FIELD, 2); continue; }
          if (e == "grad(T)[z]") { leaf(FBranch::FIELD, 3); continue; }
          if (e == "T_t") { leaf(FBranch::FIELD, 4); continue; }
          if (e == "x") { leaf(FBranch::SCALARF, 0); continue; }
          if (e == "y") { leaf(FBranch::SCALARF, 1); continue; }
          if (e == "z") { leaf(FBranch::SCALARF, 2); continue; }
          if (e == "u" || e == "v" || e == "w") { leaf(FBranch::SCALARF, 3); continue; }
          if (e == "k0" && have_k0) { leaf(FBranch::DATA, 0); continue; }
          if (e == "src" && have_src) { leaf(FBranch::DATA, 1); continue; }
          const int fi = find(e);
          if (fi >= 0) {
            b.decomposed = true;
            b.kind = FBranch::FUNC;
            b.index = fi;
            continue;
          }
          if (f_is_scalar(e)) {
            leaf(FBranch::NUMBER, -1);
            try {
              br[k].cval = std::stod(e);
            } catch (...) {
              ffail("not a number: " + e);
            }
            continue;
          }
          if (e == "t") { leaf(FBranch::TIME, -1); continue; }
          if (e == "pi") { leaf(FBranch::PI, -1); br[k].cval = 3.141592653589793238463; continue; }
          if (e == "h") ffail("'h' has no value at volume points");
          if (f_is_operator(br, k)) continue;
          const size_t n0 = br.size();
          f_split(br, k);
          br[k].decomposed = true;
          if (n0 == br.size()) ffail("not able to decompose or find: " + br[k].expr);
          // a piece equal to its branch (one unknown character) would be split
          // again without end: nothing resolves it
          if (br.size() == n0 + 1 && br.back().expr == br[k].expr)
            ffail("not able to decompose or find: " + br[k].expr);
          if (br.size() > 100000) ffail("expression too large: " + trees[ti].name);
        }
        done = true;
        for (const auto& b2 : trees[ti].b)
          if (!b2.leaf && !b2.decomposed) done = false;
      }
      if (!done) ffail("not able to decompose " + trees[ti].name);
    }
    // data types (checkDepDataType), with cycle detection
    for (size_t ti = 0; ti < trees.size(); ++ti)
      for (size_t j = 0; j < trees[ti].b.size(); ++j) {
        bool c = true, v = false, a = false;
        std::vector<std::vector<char>> busy(trees.size());
        for (size_t t2 = 0; t2 < trees.size(); ++t2) busy[t2].assign(trees[t2].b.size(), 0);
        dtype((int)ti, (int)j, c, v, a, busy);
        trees[ti].b[j].is_const = c;
        trees[ti].b[j].is_view = v;
        trees[ti].b[j].is_ad = a;
      }
  }

  void dtype(int t, int j, bool& c, bool& v, bool& a, std::vector<std::vector<char>>& busy) const {
    if (busy[t][j]) ffail("cyclic graph in: " + trees[t].b[j].expr);
    busy[t][j] = 1;
    const FBranch& b = trees[t].b[j];
    if (b.leaf) {
      if (!(b.kind == FBranch::NUMBER || b.kind == FBranch::PI)) c = false;
      if (b.kind == FBranch::FIELD || b.kind == FBranch::SCALARF || b.kind == FBranch::DATA) v = true;
      if (b.kind == FBranch::FIELD) a = true;
    } else if (b.kind == FBranch::FUNC) {
      dtype(b.index, 0, c, v, a, busy);
    } else {
      for (int d : b.deps) dtype(t, d, c, v, a, busy);
    }
    busy[t][j] = 0;
  }

  // FunctionManager::evaluate on one element
  template <class E>
  FVal<E> eval(int t, int j, const FElem<E>& el) const {
    const FBranch& b = trees[t].b[j];
    FVal<E> out;
    out.view = b.is_view;
    if (b.leaf) {
      if (b.kind == FBranch::NUMBER || b.kind == FBranch::PI) { out.s = E(b.cval); return out; }
      if (b.kind == FBranch::TIME) { out.s = E(el.time); return out; }
      out.v.assign(el.nq, E(0.0));
      for (int q = 0; q < el.nq; ++q) {
        switch (b.kind) {
          case FBranch::FIELD:
            if (b.index == 0) out.v[q] = el.T[q];
            else if (b.index <= 3) out.v[q] = (b.index - 1 < el.dim) ? el.dT[q * el.dim + b.index - 1] : E(0.0);
            else out.v[q] = el.Tt ? el.Tt[q] : E(0.0);
            break;
          case FBranch::SCALARF: out.v[q] = E(b.index < 3 ? el.x[q * 3 + b.index] : 0.0); break;
          case FBranch::DATA: out.v[q] = E(b.index == 0 ? el.k0 : el.src); break;
          default: break;
        }
      }
      return out;
    }
    if (b.kind == FBranch::FUNC) {
      FVal<E> r = eval(b.index, 0, el);
      if (r.view == out.view) return r;
      out.v.assign(el.nq, r.s);
      return out;
    }
    if (out.view) out.v.assign(el.nq, E(0.0));
    for (size_t k = 0; k < b.deps.size(); ++k) {
      const FVal<E> d = eval(t, b.deps[k], el);
      const std::string& op = b.ops[k];
      if (out.view) {
        if (d.view) vtov(out.v, d.v, op);
        else stov(out.v, d.s, op);
      } else {
        stos(out.s, d.s, op);
      }
    }
    return out;
  }

  template <class E>
  static void unary(E& data, const E& t, const std::string& op) {
    if (op == "sin") data = sin(t);
    else if (op == "cos") data = cos(t);
    else if (op == "tan") data = tan(t);
    else if (op == "exp") data = exp(t);
    else if (op == "log") data = log(t);
    else if (op == "sinh") data = sinh(t);
    else if (op == "cosh") data = cosh(t);
    else if (op == "abs") data = val(t) < 0.0 ? -t : t;
    else if (op == "sqrt") data = val(t) <= 0.0 ? E(0.0) : dsqrt(t);
    else ffail("operator " + op);
  }

  template <class E>
  static bool common(E& data, const E& t, const std::string& op) {
    if (op == "") data = t;
    else if (op == "plus") data += t;
    else if (op == "minus") data += -t;
    else if (op == "times") data = data * t;
    else if (op == "divide") data = data / t;
    else if (op == "lt") data = E(val(data) < val(t) ? 1.0 : 0.0);
    else if (op == "lte") data = E(val(data) <= val(t) ? 1.0 : 0.0);
    else if (op == "gt") data = E(val(data) > val(t) ? 1.0 : 0.0);
    else if (op == "gte") data = E(val(data) >= val(t) ? 1.0 : 0.0);
    else if (op == "max") { if (val(t) > val(data)) data = t; }
    else if (op == "min") { if (val(t) < val(data)) data = t; }
    else if (op == "mean") data = 0.5 * data + 0.5 * t;
    else return false;
    return true;
  }

  // evaluateOpVToV
  template <class E>
  static void vtov(std::vector<E>& data, const std::vector<E>& t, const std::string& op) {
    const size_t n = data.size();
    if (op == "emax" || op == "emin") {
      E r = t[0];
      for (size_t q = 0; q < n; ++q)
        if (op == "emax" ? val(t[q]) > val(t[0]) : val(t[q]) < val(t[0])) r = t[q];
      for (size_t q = 0; q < n; ++q) data[q] = r;
      return;
    }
    if (op == "emean") {
      const double scale = (double)n;
      E r = t[0] / scale;
      for (size_t q = 0; q < n; ++q) r += t[q] / scale;
      for (size_t q = 0; q < n; ++q) data[q] = r;
      return;
    }
    for (size_t q = 0; q < n; ++q) {
      if (common(data[q], t[q], op)) continue;
      if (op == "power") data[q] = pow(data[q], t[q]);
      else unary(data[q], t[q], op);
    }
  }

  // evaluateOpSToV
  template <class E>
  static void stov(std::vector<E>& data, const E& t, const std::string& op) {
    const size_t n = data.size();
    if (op == "emax" || op == "emin" || op == "emean") {
      for (size_t q = 0; q < n; ++q) data[q] = t;
      return;
    }
    if (op == "power") {
      const double ps = val(t), rd = std::round(ps);
      if (rd >= 2.0 && rd <= 4.0 && std::fabs(ps - rd) < 1.0e-12) {
        for (size_t q = 0; q < n; ++q) {
          const E base = data[q];
          E acc = base;
          for (int jj = 1; jj < (int)rd; ++jj) acc = acc * base;
          data[q] = acc;
        }
      } else {
        for (size_t q = 0; q < n; ++q) data[q] = pow(data[q], E(ps));
      }
      return;
    }
    for (size_t q = 0; q < n; ++q) {
      if (common(data[q], t, op)) continue;
      unary(data[q], t, op);
    }
  }

  // evaluateOpSToS
  template <class E>
  static void stos(E& data, const E& t, const std::string& op) {
    if (op == "emax" || op == "emin" || op == "emean") { data = t; return; }
    if (op == "power") { data = pow(data, t); return; }
    if (common(data, t, op)) return;
    unary(data, t, op);
  }

  // values of function `name` on one element (a view, or the scalar broadcast)
  template <class E>
  std::vector<E> values(const std::string& name, const FElem<E>& el) const {
    const int t = find(name);
    if (t < 0) ffail("no function " + name);
    FVal<E> r = eval(t, 0, el);
    if (r.view) return r.v;
    return std::vector<E>(el.nq, r.s);
  }
};

}  // namespace orc
