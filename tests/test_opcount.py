"""Operation counts of the device formulation (geon_amd/opcount).

The counter compiles the kernels' per-ray code for the host with a counting
scalar in place of double. The committed op_counts.json (used for the
roofline in bench.py) must be what the current sources produce, and the
Minkowski step, whose Christoffel symbols vanish, must cost exactly the RK4
combination: hh = 0.5h (1), stage 2 inputs 4·(2+2), stages 3 and 4 inputs
2·4·(2+2+2+2), h/6 (1), k4 accumulation 4·2 and the update 8·2 = 106.
"""
import json
import os

import pytest

from geon_amd import build


@pytest.fixture(scope="module")
def counts():
    return json.loads(build.run_opcount())


def test_committed_counts_are_current(counts):
    with open(build.OPCOUNT_JSON) as f:
        committed = json.load(f)
    assert committed == counts


def test_minkowski_step_is_the_rk4_combination(counts):
    assert counts["ray_step"]["minkowski"]["ops"] == 1 + 16 + 2 * 32 + 1 + 8 + 16


def test_counts_scale_with_metric_work(counts):
    rs = counts["ray_step"]
    # analytic Γ < generic FD Γ with one sine per shift < FD Γ with tanh walls
    assert rs["minkowski"]["ops"] < rs["schwarzschild"]["ops"] < rs["morris_thorne"]["ops"] \
        < rs["alcubierre_wall"]["ops"]
    # one sincos per stage: Schwarzschild for Γ, Morris-Thorne for sin θ and,
    # by angle addition, sin(θ ± h)
    assert rs["schwarzschild"]["sin"] == 4 and rs["schwarzschild"]["cos"] == 4
    assert rs["morris_thorne"]["sin"] == 4 and rs["morris_thorne"]["cos"] == 4
    # beyond the Alcubierre wall Γ is exactly zero: the step is the combination
    # plus the r_s test (3 mul, 2 add, 1 sqrt, 1 sub per stage)
    assert rs["alcubierre_flat"]["ops"] == 106 + 4 * 7
