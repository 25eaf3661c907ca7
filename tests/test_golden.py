"""The oracle against the committed golden vectors (tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "golden_small.npz")
CASES = ("tiny", "c1small", "sigma3")


def load_case(name):
    z = np.load(GOLDEN)
    return {k.split("/", 1)[1]: z[k] for k in z.files if k.startswith(name + "/")}


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_golden(oracle, name):
    g = load_case(name)
    Ke = oracle.hm_forward(g["logK"], g["neighbors"])
    assert rel(Ke, g["Ke"]) < 1e-14
    h, pb = oracle.steadystate_rrule(Ke, g["neighbors"], g["aol"], g["dirichletnodes"],
                                     g["dirichleths"], g["Qs"])
    assert rel(h, g["h"]) < 1e-10
    _, dK, _, _, _, ddh, dQ = pb(g["delta"])
    assert rel(dK, g["dK"]) < 1e-9
    assert rel(ddh, g["ddh"]) < 1e-9
    assert rel(dQ, g["dQ"]) < 1e-9
    dlogK = oracle.hm_pullback(g["logK"], g["neighbors"], dK)
    assert rel(dlogK, g["dlogK"]) < 1e-9
