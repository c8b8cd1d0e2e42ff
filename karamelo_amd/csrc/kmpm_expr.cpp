// kmpm_expr.cpp -- the input script's arithmetic for fix values that
// depend on time (fix velocity_nodes, fix_velocity_nodes.cpp:64-116).
//
// The reference keeps such a value as a Var: an equation string that
// Input::parsev (input.cpp:374-735) parses again at every evaluation.  This
// compiles the string once, following the same rules, into a program of
// time and dt that the device evaluates each step:
//   - literals are read with stof (input.cpp:407,411), so they carry float
//     precision; "e" / "E" after a digit is the power-of-ten operator
//     (input.cpp:186, 660-673);
//   - operators apply left to right by the parser's precedence table
//     (input.cpp:150-172), including its unary-minus flag, which a function
//     call does not reset (input.cpp:721-722);
//   - an operation on two constants folds at parse time (var.cpp:68-226);
//     a constant inside a time-dependent expression is written into the
//     equation as "%.15f" (var.cpp:24-29) and read back with stof at each
//     evaluation, so the program holds stof("%.15f" of the constant);
//   - a wholly constant expression keeps the double value of its fold.
// The previous value of a fix replaces the text "time" by "time - dt"
// before parsing (fix_velocity_nodes.cpp:69-77), without parentheses, as
// the reference does.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/kmpm.h"

namespace {

struct Item {                  // a Var: constant, or a program of time and dt
  bool constant = true;
  double value = 0;
  std::vector<int> op;
  std::vector<double> c;
};

struct Fail {};

double read_float(const std::string &s) { return (double)std::strtof(s.c_str(), nullptr); }

// A constant as the re-parsed equation holds it: stof of its "%.15f" text.
double reparsed(double v) {
  char b[512];
  std::snprintf(b, sizeof(b), "%.15f", std::fabs(v));
  const double r = read_float(b);
  return std::signbit(v) ? -r : r;
}

void emit(const Item &x, Item &out) {
  if (x.constant) {
    out.op.push_back(KMPM_OP_CONST);
    out.c.push_back(reparsed(x.value));
  } else {
    out.op.insert(out.op.end(), x.op.begin(), x.op.end());
    out.c.insert(out.c.end(), x.c.begin(), x.c.end());
  }
}

Item leaf(int op) {
  Item r;
  r.constant = false;
  r.op.push_back(op);
  r.c.push_back(0.0);
  return r;
}

Item constant(double v) {
  Item r;
  r.value = v;
  return r;
}

Item unary(const Item &x, int op) {
  if (x.constant) {
    switch (op) {
      case KMPM_OP_NEG: return constant(-x.value);
      case KMPM_OP_NOT: return constant(!x.value);
      case KMPM_OP_EXP: return constant(std::exp(x.value));
      case KMPM_OP_SQRT: return constant(std::sqrt(x.value));
      case KMPM_OP_COS: return constant(std::cos(x.value));
      case KMPM_OP_SIN: return constant(std::sin(x.value));
      case KMPM_OP_TAN: return constant(std::tan(x.value));
      case KMPM_OP_LOG: return constant(std::log(x.value));
    }
    throw Fail();
  }
  Item r;
  r.constant = false;
  emit(x, r);
  r.op.push_back(op);
  r.c.push_back(0.0);
  return r;
}

Item binary(const Item &a, int op, const Item &b) {
  if (a.constant && b.constant) {
    const double x = a.value, y = b.value;
    switch (op) {
      case KMPM_OP_ADD: return constant(x + y);
      case KMPM_OP_SUB: return constant(x - y);
      case KMPM_OP_MUL: return constant(x * y);
      case KMPM_OP_DIV: return constant(x / y);
      case KMPM_OP_POW: return constant(std::pow(x, y));
      case KMPM_OP_GT: return constant(x > y);
      case KMPM_OP_GE: return constant(x >= y);
      case KMPM_OP_LT: return constant(x < y);
      case KMPM_OP_LE: return constant(x <= y);
      case KMPM_OP_EQ: return constant(x == y);
      case KMPM_OP_NE: return constant(x != y);
      case KMPM_OP_ATAN2: return constant(std::atan2(x, y));
    }
    throw Fail();
  }
  Item r;
  r.constant = false;
  emit(a, r);
  emit(b, r);
  r.op.push_back(op);
  r.c.push_back(0.0);
  return r;
}

// Input::applyOp (input.cpp:180-212); "e" is a * 10^b (powv, var.cpp:228-236).
Item apply(const Item &a, const std::string &op, const Item &b) {
  if (op == "+") return binary(a, KMPM_OP_ADD, b);
  if (op == "-") return binary(a, KMPM_OP_SUB, b);
  if (op == "*") return binary(a, KMPM_OP_MUL, b);
  if (op == "/") return binary(a, KMPM_OP_DIV, b);
  if (op == "^" || op == "**") return binary(a, KMPM_OP_POW, b);
  if (op == "e") {
    if (!b.constant) throw Fail();           // the reference cannot evaluate pow(10, ...)
    return binary(a, KMPM_OP_MUL, constant(std::pow(10, b.value)));
  }
  if (op == ">") return binary(a, KMPM_OP_GT, b);
  if (op == "<") return binary(a, KMPM_OP_LT, b);
  if (op == ">=") return binary(a, KMPM_OP_GE, b);
  if (op == "<=") return binary(a, KMPM_OP_LE, b);
  if (op == "==") return binary(a, KMPM_OP_EQ, b);
  if (op == "!=") return binary(a, KMPM_OP_NE, b);
  throw Fail();
}

int precedence(const std::string &op) {           // input.cpp:150-172
  if (op[0] == '>' || op[0] == '<' || op[0] == '=' || op[0] == '!') return 1;
  if (op[0] == '+' || op[0] == '-') return 2;
  if (op[0] == '/') return 3;
  if (op[0] == '^') return 4;
  if (op[0] == 'e' || op[0] == 'E') return 5;
  if (op == "*") return 3;
  if (op == "**") return 4;
  return 0;
}

bool is_operator(char c) {
  return c == '+' || c == '-' || c == '*' || c == '/' || c == '^' || c == '>' || c == '<' ||
         c == '!';
}
bool is_math_char(char c) { return is_operator(c) || c == '(' || c == ')' || c == '='; }

struct Parser {
  int nvar;
  const char *const *names;
  const double *values;

  bool variable(const std::string &w, Item &out) const {
    if (w == "time") { out = leaf(KMPM_OP_TIME); return true; }
    if (w == "dt") { out = leaf(KMPM_OP_DT); return true; }
    if (w == "PI") { out = constant(M_PI); return true; }   // input.cpp:62
    for (int q = nvar - 1; q >= 0; q--)
      if (names[q] && w == names[q]) { out = constant(values[q]); return true; }
    return false;
  }

  Item function(const std::string &f, const std::string &arg) const {
    static const struct { const char *name; int op; } fns[] = {
        {"exp", KMPM_OP_EXP}, {"sqrt", KMPM_OP_SQRT}, {"cos", KMPM_OP_COS},
        {"sin", KMPM_OP_SIN}, {"tan", KMPM_OP_TAN}, {"log", KMPM_OP_LOG}};
    for (const auto &q : fns)
      if (f == q.name) return unary(parse(arg), q.op);
    if (f == "atan2") {                          // input.cpp:258-263
      const size_t k = arg.find(',');
      if (k == std::string::npos || arg.find(',', k + 1) != std::string::npos) throw Fail();
      return binary(parse(arg.substr(0, k)), KMPM_OP_ATAN2, parse(arg.substr(k + 1)));
    }
    throw Fail();
  }

  // Applies the top operator the way the ')' and end-of-string loops do
  // (input.cpp:436-480, 690-716): with one value left, a unary -, ! or +.
  static void reduce(std::vector<Item> &values, std::vector<std::string> &ops) {
    if (values.empty()) throw Fail();
    if (values.size() == 1) {
      const std::string op = ops.back();
      ops.pop_back();
      if (op == "-") values.back() = unary(values.back(), KMPM_OP_NEG);
      else if (op == "!") values.back() = unary(values.back(), KMPM_OP_NOT);
      else if (op != "+") throw Fail();
      return;
    }
    Item b = values.back();
    values.pop_back();
    Item a = values.back();
    values.pop_back();
    const std::string op = ops.back();
    ops.pop_back();
    values.push_back(apply(a, op, b));
  }

  // Input::parsev (input.cpp:374-735), arithmetic part.
  Item parse(std::string str) const {
    std::string s;
    for (char ch : str)
      if (ch != ' ') s.push_back(ch);
    std::vector<Item> values;
    std::vector<std::string> ops;
    bool negative = false;
    const int L = (int)s.size();
    for (int i = 0; i < L; i++) {
      const char ch = s[i];
      if (isdigit((unsigned char)ch)) {
        int j = 1;
        for (; j < L - i; j++)
          if (!isdigit((unsigned char)s[i + j]) && s[i + j] != '.') break;
        const double v = read_float(s.substr(i, j));
        i += j - 1;
        values.push_back(constant(negative ? (-1) * v : v));
        negative = false;
      } else if (ch == '(') {
        ops.push_back("(");
        if (i + 1 < L && s[i + 1] == '-' && !(i + 2 < L && s[i + 2] == '(')) {
          negative = true;
          i++;
        }
      } else if (ch == ')') {
        if (ops.empty() && values.size() < 2) throw Fail();
        while (!ops.empty() && ops.back() != "(") reduce(values, ops);
        if (ops.empty()) throw Fail();
        ops.pop_back();
      } else if (is_operator(ch) || (ch == '=' && i + 1 < L && s[i + 1] == '=')) {
        std::string op(1, ch);
        if (values.empty() && !(i + 1 < L && s[i + 1] == '(')) {
          if (op == "-") {
            negative = true;
            continue;
          }
        }
        if (i + 1 >= L) throw Fail();
        if (s[i + 1] == '*' || s[i + 1] == '=') {
          op.push_back(s[i + 1]);
          i++;
        } else if (is_operator(s[i + 1]) && s[i + 1] != '-') {
          throw Fail();
        }
        while (!ops.empty() && precedence(ops.back()) >= precedence(op)) reduce(values, ops);
        ops.push_back(op);
        if (i + 1 < L && s[i + 1] == '-') {
          negative = true;
          i++;
        }
      } else {
        int j = 1;
        for (; j < L - i; j++)
          if (is_math_char(s[i + j])) break;
        const std::string word = s.substr(i, j);
        i += j - 1;
        if ((word == "E" || word == "e") && !values.empty() && i >= 1 &&
            isdigit((unsigned char)s[i - 1]) && i + 1 < L && (s[i + 1] == '+' || s[i + 1] == '-')) {
          ops.push_back("e");
          if (s[i + 1] == '-') {
            negative = true;
            i++;
          }
          if (i + 1 < L && s[i + 1] == '+') i++;
          continue;
        }
        Item var;
        if (variable(word, var)) {
          values.push_back(negative ? unary(var, KMPM_OP_NEG) : var);
          negative = false;
        } else if (i + 1 < L && s[i + 1] == '(') {
          i += 2;
          int k = 0, nl = 0, nr = 0;
          while (s[i + k] != ')' || nl != nr) {
            if (s[i + k] == '(') nl++;
            if (s[i + k] == ')') nr++;
            k++;
            if (i + k >= L) throw Fail();
          }
          const Item f = function(word, s.substr(i, k));
          i += k;
          values.push_back(negative ? unary(f, KMPM_OP_NEG) : f);   // the flag stays set
        } else {
          throw Fail();
        }
      }
    }
    while (!ops.empty()) reduce(values, ops);
    if (values.empty()) return constant(-1);        // input.cpp:719-724
    return values.back();
  }
};

}  // namespace

extern "C" int kmpm_expr_compile(const char *text, int previous, int nvar,
                                 const char *const *names, const double *values,
                                 kmpm_expr *out) {
  if (!text || !out || nvar < 0 || (nvar > 0 && (!names || !values))) return KMPM_ERR_ARG;
  std::memset(out, 0, sizeof(*out));
  std::string s(text);
  if (previous) {                 // fix_velocity_nodes.cpp:69-77
    const std::string from = "time", to = "time - dt";
    for (size_t p = s.find(from); p != std::string::npos; p = s.find(from, p + to.size()))
      s.replace(p, from.size(), to);
  }
  Item r;
  try {
    r = Parser{nvar, names, values}.parse(s);
  } catch (Fail &) {
    return KMPM_ERR_ARG;
  }
  if (r.constant) {
    out->n = 1;
    out->op[0] = KMPM_OP_CONST;
    out->c[0] = r.value;
    return KMPM_OK;
  }
  if (r.op.size() > (size_t)KMPM_EXPR_MAX) return KMPM_ERR_CAPACITY;
  // the evaluation stack of the device holds KMPM_EXPR_STACK values
  int depth = 0, deepest = 0;
  for (size_t q = 0; q < r.op.size(); q++) {
    const int o = r.op[q];
    if (o == KMPM_OP_CONST || o == KMPM_OP_TIME || o == KMPM_OP_DT) depth++;
    else if (o >= KMPM_OP_ADD && o <= KMPM_OP_NE) depth--;
    else if (o == KMPM_OP_ATAN2) depth--;
    if (depth < 1) return KMPM_ERR_ARG;
    deepest = depth > deepest ? depth : deepest;
    out->op[q] = o;
    out->c[q] = r.c[q];
  }
  if (depth != 1 || deepest > KMPM_EXPR_STACK) return KMPM_ERR_CAPACITY;
  out->n = (int)r.op.size();
  return KMPM_OK;
}
