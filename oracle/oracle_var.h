/*
 * oracle_var.h -- restatement of the reference's expression variables, for
 * the oracle (TEST INFRASTRUCTURE ONLY).
 *
 * Var (var.cpp:24-309) holds an equation string, a value and a constant
 * flag; Input::parsev (input.cpp:374-735) evaluates a string with a value
 * stack and an operator stack.  Operations on two constants give a new
 * constant whose equation is "%.15f" of its value (var.cpp:24-29,
 * Var::Var(double)); any other operation builds the equation "(" a op b ")"
 * from the operands' equations.  Var::result re-parses a non-constant
 * equation at each call (var.cpp:44-59), so a fix value is evaluated every
 * step by parsing its equation again with the current time and dt.  This
 * file restates exactly that, string for string: it is independent of the
 * product, which compiles the same text once into a device program.
 * Supported: the arithmetic, comparison and function subset a fix value
 * uses (+ - * / ^ ** e E > >= < <= == != !, exp sqrt cos sin tan log atan2)
 * and the variables time, dt, PI and constant user variables.
 */
#ifndef ORACLE_VAR_H
#define ORACLE_VAR_H

#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define OV_EQ 1024
#define OV_STACK 64

typedef struct {
  int constant;
  double value;
  char eq[OV_EQ];
} ovar;

typedef struct {
  double time, dt;
  int nvar;
  const char *const *names;
  const double *values;
  int error;
} ovenv;

static ovar ov_const(double v) {                 /* Var::Var(double), var.cpp:31-36 */
  ovar r;
  r.constant = 1;
  r.value = v;
  snprintf(r.eq, OV_EQ, "%.15f", v);
  return r;
}

static ovar ov_named(const char *name, double v) {  /* Var(eq, v), c = false */
  ovar r;
  r.constant = 0;
  r.value = v;
  snprintf(r.eq, OV_EQ, "%s", name);
  return r;
}

/* Var::operator<op> (var.cpp:68-226): fold two constants, else "(a op b)" */
static ovar ov_bin(ovar a, const char *op, ovar b, double v) {
  if (a.constant && b.constant) return ov_const(v);
  ovar r;
  r.constant = 0;
  r.value = v;
  snprintf(r.eq, OV_EQ, "(%s%s%s)", a.eq, op, b.eq);
  return r;
}

static ovar ov_neg(ovar a) {                     /* Var::operator-() */
  if (a.constant) return ov_const(-a.value);
  ovar r;
  r.constant = 0;
  r.value = -a.value;
  snprintf(r.eq, OV_EQ, "(-%s)", a.eq);
  return r;
}

static ovar ov_not(ovar a) {                     /* Var::operator!() */
  if (a.constant) return ov_const(!a.value);
  ovar r;
  r.constant = 0;
  r.value = !a.value;
  snprintf(r.eq, OV_EQ, "(!%s)", a.eq);
  return r;
}

static ovar ov_fn(const char *f, ovar x, double v) {  /* expv ... logv, var.cpp:238-299 */
  if (x.constant) return ov_const(v);
  ovar r;
  r.constant = 0;
  r.value = v;
  snprintf(r.eq, OV_EQ, "%s(%s)", f, x.eq);
  return r;
}

/* Input::applyOp (input.cpp:180-212) */
static ovar ov_apply(ovar a, const char *op, ovar b, ovenv *env) {
  double x = a.value, y = b.value;
  if (!strcmp(op, "+")) return ov_bin(a, "+", b, x + y);
  if (!strcmp(op, "-")) return ov_bin(a, "-", b, x - y);
  if (!strcmp(op, "*")) return ov_bin(a, "*", b, x * y);
  if (!strcmp(op, "/")) return ov_bin(a, "/", b, x / y);
  if (!strcmp(op, "^") || !strcmp(op, "**")) return ov_bin(a, "^", b, pow(x, y));
  if (!strcmp(op, "e")) {                        /* a * powv(10, b), var.cpp:228-236 */
    if (!b.constant) { env->error = 1; return a; }
    ovar p = ov_const(pow(10, y));
    return ov_bin(a, "*", p, x * p.value);
  }
  if (!strcmp(op, ">")) return ov_bin(a, ">", b, x > y);
  if (!strcmp(op, "<")) return ov_bin(a, "<", b, x < y);
  if (!strcmp(op, ">=")) return ov_bin(a, ">=", b, x >= y);
  if (!strcmp(op, "<=")) return ov_bin(a, "<=", b, x <= y);
  if (!strcmp(op, "==")) return ov_bin(a, "==", b, x == y);
  if (!strcmp(op, "!=")) return ov_bin(a, "!=", b, x != y);
  env->error = 1;
  return a;
}

static int ov_prec(const char *op) {             /* Input::precedence, input.cpp:150-172 */
  if (op[0] == '>' || op[0] == '<' || op[0] == '=' || op[0] == '!') return 1;
  if (op[0] == '+' || op[0] == '-') return 2;
  if (op[0] == '/') return 3;
  if (op[0] == '^') return 4;
  if (op[0] == 'e' || op[0] == 'E') return 5;
  if (!strcmp(op, "*")) return 3;
  if (!strcmp(op, "**")) return 4;
  return 0;
}

static int ov_isop(char c) {
  return c == '+' || c == '-' || c == '*' || c == '/' || c == '^' || c == '>' || c == '<' ||
         c == '!';
}
static int ov_ismath(char c) { return ov_isop(c) || c == '(' || c == ')' || c == '='; }

static ovar ov_parse(const char *text, ovenv *env);

/* Input::evaluate_function, the functions a value can use (input.cpp:244-263) */
static ovar ov_function(const char *f, const char *arg, ovenv *env) {
  if (!strcmp(f, "atan2")) {
    const char *comma = strchr(arg, ',');
    if (!comma || strchr(comma + 1, ',')) { env->error = 1; return ov_const(0); }
    char a0[OV_EQ];
    snprintf(a0, OV_EQ, "%.*s", (int)(comma - arg), arg);
    ovar x = ov_parse(a0, env), y = ov_parse(comma + 1, env);
    if (x.constant && y.constant) return ov_const(atan2(x.value, y.value));
    ovar r;                                      /* atan2v, var.cpp:301-309 */
    r.constant = 0;
    r.value = atan2(x.value, y.value);
    snprintf(r.eq, OV_EQ, "atan2(%s, %s)", x.eq, y.eq);
    return r;
  }
  ovar x = ov_parse(arg, env);
  if (!strcmp(f, "exp")) return ov_fn("exp", x, exp(x.value));
  if (!strcmp(f, "sqrt")) return ov_fn("sqrt", x, sqrt(x.value));
  if (!strcmp(f, "cos")) return ov_fn("cos", x, cos(x.value));
  if (!strcmp(f, "sin")) return ov_fn("sin", x, sin(x.value));
  if (!strcmp(f, "tan")) return ov_fn("tan", x, tan(x.value));
  if (!strcmp(f, "log")) return ov_fn("log", x, log(x.value));
  env->error = 1;
  return x;
}

static int ov_lookup(const char *w, ovenv *env, ovar *out) {
  if (!strcmp(w, "time")) { *out = ov_named("time", env->time); return 1; }
  if (!strcmp(w, "dt")) { *out = ov_named("dt", env->dt); return 1; }
  if (!strcmp(w, "PI")) { *out = ov_const(3.14159265358979323846); return 1; }  /* M_PI, input.cpp:62 */
  for (int q = env->nvar - 1; q >= 0; q--)
    if (env->names[q] && !strcmp(w, env->names[q])) { *out = ov_const(env->values[q]); return 1; }
  return 0;
}

typedef struct {
  ovar v[OV_STACK];
  int nv;
  char op[OV_STACK][4];
  int nop;
} ovstacks;

/* the closing loops of input.cpp:436-480 and 690-716 */
static void ov_reduce(ovstacks *st, ovenv *env) {
  if (st->nv < 1 || st->nop < 1) { env->error = 1; st->nop = 0; return; }
  const char *op = st->op[--st->nop];
  if (st->nv == 1) {
    if (!strcmp(op, "-")) st->v[0] = ov_neg(st->v[0]);
    else if (!strcmp(op, "!")) st->v[0] = ov_not(st->v[0]);
    else if (strcmp(op, "+")) env->error = 1;
    return;
  }
  ovar b = st->v[--st->nv], a = st->v[--st->nv];
  st->v[st->nv++] = ov_apply(a, op, b, env);
}

/* Input::parsev (input.cpp:374-735), arithmetic part */
static ovar ov_parse(const char *text, ovenv *env) {
  char s[OV_EQ];
  int L = 0;
  for (const char *p = text; *p && L < OV_EQ - 1; p++)
    if (*p != ' ') s[L++] = *p;
  s[L] = 0;
  ovstacks st;
  st.nv = st.nop = 0;
  int negative = 0;
  for (int i = 0; i < L && !env->error; i++) {
    char c = s[i];
    if (st.nv >= OV_STACK - 1 || st.nop >= OV_STACK - 1) { env->error = 1; break; }
    if (isdigit((unsigned char)c)) {
      int j = 1;
      while (j < L - i && (isdigit((unsigned char)s[i + j]) || s[i + j] == '.')) j++;
      char num[OV_EQ];
      snprintf(num, OV_EQ, "%.*s", j, s + i);
      i += j - 1;
      double v = (double)strtof(num, NULL);      /* stof, input.cpp:407,411 */
      st.v[st.nv++] = ov_const(negative ? (-1) * v : v);
      negative = 0;
    } else if (c == '(') {
      strcpy(st.op[st.nop++], "(");
      if (i + 1 < L && s[i + 1] == '-' && !(i + 2 < L && s[i + 2] == '(')) { negative = 1; i++; }
    } else if (c == ')') {
      if (st.nop == 0 && st.nv < 2) { env->error = 1; break; }
      while (st.nop > 0 && strcmp(st.op[st.nop - 1], "(") && !env->error) ov_reduce(&st, env);
      if (st.nop == 0) { env->error = 1; break; }
      st.nop--;
    } else if (ov_isop(c) || (c == '=' && i + 1 < L && s[i + 1] == '=')) {
      char op[4] = {c, 0, 0, 0};
      if (st.nv == 0 && !(i + 1 < L && s[i + 1] == '(') && c == '-') { negative = 1; continue; }
      if (i + 1 >= L) { env->error = 1; break; }
      if (s[i + 1] == '*' || s[i + 1] == '=') { op[1] = s[i + 1]; i++; }
      else if (ov_isop(s[i + 1]) && s[i + 1] != '-') { env->error = 1; break; }
      while (st.nop > 0 && ov_prec(st.op[st.nop - 1]) >= ov_prec(op) && !env->error)
        ov_reduce(&st, env);
      strcpy(st.op[st.nop++], op);
      if (i + 1 < L && s[i + 1] == '-') { negative = 1; i++; }
    } else {
      int j = 1;
      while (j < L - i && !ov_ismath(s[i + j])) j++;
      char w[OV_EQ];
      snprintf(w, OV_EQ, "%.*s", j, s + i);
      i += j - 1;
      if ((!strcmp(w, "E") || !strcmp(w, "e")) && st.nv > 0 && i >= 1 &&
          isdigit((unsigned char)s[i - 1]) && i + 1 < L && (s[i + 1] == '+' || s[i + 1] == '-')) {
        strcpy(st.op[st.nop++], "e");
        if (s[i + 1] == '-') { negative = 1; i++; }
        if (i + 1 < L && s[i + 1] == '+') i++;
        continue;
      }
      ovar v;
      if (ov_lookup(w, env, &v)) {
        st.v[st.nv++] = negative ? ov_neg(v) : v;
        negative = 0;
      } else if (i + 1 < L && s[i + 1] == '(') {
        i += 2;
        int k = 0, nl = 0, nr = 0;
        while (s[i + k] != ')' || nl != nr) {
          if (s[i + k] == '(') nl++;
          if (s[i + k] == ')') nr++;
          k++;
          if (i + k >= L) { env->error = 1; break; }
        }
        if (env->error) break;
        char arg[OV_EQ];
        snprintf(arg, OV_EQ, "%.*s", k, s + i);
        i += k;
        ovar f = ov_function(w, arg, env);
        st.v[st.nv++] = negative ? ov_neg(f) : f;   /* the flag is not reset, input.cpp:721-722 */
      } else {
        env->error = 1;
      }
    }
  }
  while (st.nop > 0 && !env->error) ov_reduce(&st, env);
  if (st.nv == 0) return ov_const(-1);
  return st.v[st.nv - 1];
}

/* The fix's text with "time" replaced by "time - dt" (fix_velocity_nodes.cpp:69-77). */
static void ov_previous_text(const char *text, char *out, size_t n) {
  size_t o = 0;
  for (const char *p = text; *p && o + 10 < n;) {
    if (!strncmp(p, "time", 4)) {
      memcpy(out + o, "time - dt", 9);
      o += 9;
      p += 4;
    } else {
      out[o++] = *p++;
    }
  }
  out[o] = 0;
}

/* Var::result (var.cpp:52-59): a constant's value, else its equation parsed again. */
static double ov_result(const ovar *v, double time, double dt, int *error) {
  if (v->constant) return v->value;
  ovenv env = {time, dt, 0, NULL, NULL, 0};
  ovar r = ov_parse(v->eq, &env);
  if (env.error) *error = 1;
  return r.value;
}

#endif
