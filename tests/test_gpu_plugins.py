"""GPU parity of the constitutive plugins the headline workloads leave
unexercised, against the oracle, through the C ABI:

- `eos linear` + `strength linear` (eos_linear.cpp:71-74,
  strength_linear.cpp:59-71) on the tensile bar, USL cubic and MUSL linear;
- the shock EOS's artificial viscosity (eos_shock.cpp:99-133 with Q1, Q2
  nonzero, applied where tr D < 0) on the TLMPM Taylor impact, and on the
  same impact with ULMPM.

Bar: 1e-10 relative L2 on every particle field after N steps.
"""
import numpy as np
import pytest

import oracle_lib as O
from karamelo_amd import problems as P
from test_gpu_parity import run_and_compare, say

pytestmark = pytest.mark.gpu


def linear_steel():
    """Steel as `eos linear` (rho0, K) + `strength linear` (G): E = 200 GPa,
    nu = 0.3."""
    E, nu, rho = 200e9, P.lit(0.3), 7850.0
    K = E / (3 * (1 - 2 * nu))
    G = E / (2 * (1 + nu))
    return P.eos_strength_material("linear", (rho, K), "linear", (G,))


@pytest.mark.parametrize("scheme,shape", [("usl", "cubic-spline"), ("musl", "linear")])
def test_linear_eos_and_strength_tensile_bar(scheme, shape):
    say("eos linear + strength linear tensile bar, %s %s, 200 steps" % (scheme, shape))
    prob = P.tensile_bar(h=6e-4, v_pull=40.0, scheme=scheme, shape=shape,
                         material=linear_steel())
    w, po = run_and_compare(prob, 200, 100)
    # the bar is stretched: a nonzero deviatoric stress and a pressure of
    # both signs (the ends pull, the waves reflect)
    s = po["sigma"].reshape(-1, 9)
    assert np.abs(s[:, 1]).max() > 0 or np.abs(s[:, 0] - s[:, 4]).max() > 0


@pytest.mark.parametrize("method", ["tlmpm", "ulmpm"])
def test_shock_artificial_viscosity_taylor_impact(method):
    say("shock EOS with Q1 = 1.5, Q2 = 0.06: %s Taylor impact, 60 steps" % method)
    prob = P.taylor_bar(h=8e-4, material=P.copper_jc(q=(1.5, 0.06)), method=method)
    w, po = run_and_compare(prob, 60, 20)
    # the impact compresses the bar's end, where the viscosity acts: the
    # oracle's stress without it differs at O(1e-2)
    s = po["sigma"].reshape(-1, 9)
    assert (s[:, 0] + s[:, 4] + s[:, 8]).min() < 0
    o = O.oracle_for(P.taylor_bar(h=8e-4, method=method))
    assert o.step(60) == 0
    s0 = o.download()["sigma"]
    assert np.linalg.norm(po["sigma"] - s0) > 1e-2 * np.linalg.norm(s0)
