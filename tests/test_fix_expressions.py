"""Fix values that depend on time (fix velocity_nodes with expressions of the
input script), on the CPU.

- The product compiles a value's text once into a program of time and dt
  (kmpm_expr_compile); the oracle restates the reference's Var, which parses
  its equation string again at each evaluation (var.cpp:24-309,
  input.cpp:374-735).  The two must give the same double at every (time, dt).
- Known answers pin the oracle: the previous value replaces the text "time"
  by "time - dt" without parentheses (fix_velocity_nodes.cpp:69-77), so
  "10*time" gives 10 t - dt but "10*(time)" gives 10 (t - dt); literals are
  read with stof; a constant inside a time-dependent equation is re-read from
  its "%.15f" text.
- On a grid: v_update = value(t), v = value(t - dt) on the group's nodes, and
  the reaction force sum m (value - v_update) / dt (fix_velocity_nodes.cpp:
  130-225).
"""
import math

import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import native as N
from karamelo_amd import problems as P


def f32(x):
    return float(np.float32(x))


def rpn_eval(e, t, dt):
    """The device's evaluation of a compiled program (kmpm_kernels.h
    expr_eval), in Python doubles."""
    st = []
    bin_ops = {N_OP["ADD"]: lambda x, y: x + y, N_OP["SUB"]: lambda x, y: x - y,
               N_OP["MUL"]: lambda x, y: x * y, N_OP["DIV"]: lambda x, y: x / y,
               N_OP["POW"]: lambda x, y: math.pow(x, y), N_OP["GT"]: lambda x, y: float(x > y),
               N_OP["GE"]: lambda x, y: float(x >= y), N_OP["LT"]: lambda x, y: float(x < y),
               N_OP["LE"]: lambda x, y: float(x <= y), N_OP["EQ"]: lambda x, y: float(x == y),
               N_OP["NE"]: lambda x, y: float(x != y), N_OP["ATAN2"]: math.atan2}
    un_ops = {N_OP["NEG"]: lambda x: -x, N_OP["NOT"]: lambda x: float(not x),
              N_OP["EXP"]: math.exp, N_OP["SQRT"]: math.sqrt, N_OP["COS"]: math.cos,
              N_OP["SIN"]: math.sin, N_OP["TAN"]: math.tan, N_OP["LOG"]: math.log}
    for q in range(e.n):
        o = e.op[q]
        if o == N_OP["CONST"]:
            st.append(e.c[q])
        elif o == N_OP["TIME"]:
            st.append(t)
        elif o == N_OP["DT"]:
            st.append(dt)
        elif o in bin_ops:
            y = st.pop()
            st[-1] = bin_ops[o](st[-1], y)
        else:
            st[-1] = un_ops[o](st[-1])
    assert len(st) == 1
    return st[0]


N_OP = {k: i for i, k in enumerate(
    ["CONST", "TIME", "DT", "ADD", "SUB", "MUL", "DIV", "POW", "GT", "GE", "LT", "LE", "EQ", "NE",
     "ATAN2", "NEG", "NOT", "EXP", "SQRT", "COS", "SIN", "TAN", "LOG"])}

EXPRESSIONS = [
    "10*time", "10*(time)", "v*time", "time*v", "-v*time", "0.1", "-0.1*time", "2*3*time",
    "1e-3*time", "2.5e+2*time/3", "v*time/1e-5", "exp(-time)+2", "-exp(time)+2",
    "sin(2*PI*time/0.01)*v", "v*(1-cos(PI*time/2e-5))/2", "sqrt(time+1)", "log(1+time*1e+4)",
    "atan2(time, 1e-5)", "(time>1e-5)*v", "(time<=2e-5)*v*time", "time^2*3", "time**0.5",
    "v - -time", "(-time)*v", "(-v)*(time-1e-6)", "!time", "v", "2*PI", "v*tan(time)",
    "time==0", "(time!=0)+time", "v*time+w*time*time", "-(time)", "3-2-time",
]
VARS = {"v": 10.0, "w": -3.7e4}


@pytest.mark.parametrize("text", EXPRESSIONS)
def test_compiled_program_equals_the_reference_var(text):
    for previous in (False, True):
        e = N.expr_compile(text, previous, VARS)
        for t in (0.0, 1.25e-6, 3.3e-5, 0.125, 1.7):
            for dt in (0.0, 1e-7, 2.6e-9):
                want = O.var_eval(text, previous, t, dt, VARS)
                got = rpn_eval(e, t, dt)
                assert want is not None
                assert got == want or (math.isnan(got) and math.isnan(want)), \
                    (text, previous, t, dt, got, want)


@pytest.mark.parametrize("text", ["foo*time", "time+", "2**", "pow(2,time)", "1e3*time",
                                  "atan2(time)", "(time"])
def test_rejected_like_the_reference(text):
    with pytest.raises(N.KmpmError):
        N.expr_compile(text, False, VARS)
    assert O.var_eval(text, False, 1.0, 0.0, VARS) is None


def test_previous_value_substitutes_text_without_parentheses():
    """Known answers of the reference's "time" -> "time - dt" rewrite."""
    t, dt = 3.0e-5, 2.0e-7
    ve = lambda s, p: O.var_eval(s, p, t, dt, VARS)
    assert ve("10*time", False) == 10 * t
    assert ve("10*time", True) == 10 * t - dt                 # (10 time) - dt
    assert ve("10*(time)", True) == 10 * (t - dt)
    assert ve("time*10", True) == t - dt * 10                 # time - (dt 10)
    assert ve("v*time", True) == 10 * t - dt
    assert ve("0.1", True) == f32(0.1)                         # stof literal, no time
    assert ve("0.1*time", False) == f32(0.1) * t
    # a folded constant inside a time-dependent equation: "%.15f" then stof
    assert ve("2*PI*time", False) == f32(float("%.15f" % (2 * math.pi))) * t
    assert ve("2*PI", False) == 2 * math.pi                    # wholly constant: double
    assert ve("1e-3*time", False) == f32(0.001) * t


def _rest_particle_problem(dt=1e-3, fix_texts=("2.0", None, None)):
    cfg = P.make_config(2, (0.0, 0.0), (1.0, 1.0), 0.1, dt=dt, dt_constant=True)
    m = P.linear_material(1000.0, 1e6, 0.3)
    arrs = N.empty_particles(1)
    arrs["ptag"][0] = 1
    arrs["x"][0] = arrs["x0"][0] = [0.437, 0.529, 0.0]
    arrs["F"][0] = np.eye(3).ravel()
    arrs["mass"][0] = 2.5
    arrs["vol0"][0] = arrs["vol"][0] = 2.5e-3
    o = O.Oracle(cfg, [m])
    o.upload(arrs)
    o.set_node_mask(np.full(o.nn, 3, dtype=np.uint32))     # every node in group bit 2
    o.fix_velocity_nodes_expr(2, list(fix_texts), VARS)
    return o


def test_reaction_force_of_a_held_particle():
    """A particle at rest, every node held at vx = 2: v_update before the fix
    is 0 on every node, so the reaction is sum_I m_I 2 / dt = 2 m_p / dt
    (partition of unity of the node masses)."""
    dt = 1e-3
    o = _rest_particle_problem(dt)
    assert o.step(1) == 0
    f = o.fix_force(0)
    assert abs(f[0] - 2.0 * 2.5 / dt) <= 1e-12 * 2.0 * 2.5 / dt
    assert f[1] == 0 and f[2] == 0
    o.close()


def test_grid_holds_value_and_previous_value():
    """After step n (set_dt): the group's nodes hold v_update = value(t) and
    v = value(t - dt) with t the time at the start of the step; the first
    step of a run without set_dt sees dt = 0 (the script's dt variable)."""
    dt = 1e-3
    o = _rest_particle_problem(dt, ("v*time", "3*(time)", None))
    assert o.step(3) == 0
    _, v, vu, _ = o.grid()
    t = 2 * dt
    assert np.all(vu[:, 0] == 10 * t) and np.all(v[:, 0] == 10 * t - dt)
    assert np.all(vu[:, 1] == 3 * t) and np.all(v[:, 1] == 3 * (t - dt))
    o.close()
