"""The oracle against the golden fixtures written from the reference build
(tests/golden/make_golden.py): config 1 in full in float and double, and an
impact block on four slabs, to the bit."""
import os

import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.runner import run_scene

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STATUS = ("total_steps", "first_failure_step", "last_failure_step", "failure_rounds", "broken_bonds",
          "removed_energy_j", "last_energy_gain_step", "max_damage", "last_damage_gain_step", "exit_reason")

CASES = {
    "c1_float": lambda: scenes.config1(),
    "c1_double": lambda: scenes.config1(),
    "impact_4slabs": lambda: scenes.small_impact(substeps=150, blocks=4),
}


def load(name):
    return np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)


def check_against_golden(res, g):
    fdt = g["u"].dtype
    assert np.array_equal(res["u"].astype(fdt), g["u"]), f"max |du| {np.abs(res['u'] - g['u']).max()}"
    assert np.array_equal(res["v"].astype(fdt), g["v"])
    assert np.array_equal(np.packbits(res["alive"]), g["alive"])
    assert np.array_equal(res["mode"], g["mode"])
    assert np.array_equal(res["labels"], g["labels"])
    for f in STATUS:
        assert getattr(res["status"], f) == g[f][()], f
    assert list(res["sphere"].center) == list(g["sphere_center"])


def run_case(lib, name, **kw):
    g = load(name)
    sc = CASES[name]()
    sc.settings.direct_arithmetic = int(g["direct"])
    return run_scene(lib, sc, int(g["precision"]), **kw), g


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_reproduces_golden(oracle, name):
    res, g = run_case(oracle, name)
    check_against_golden(res, g)


def test_float_stays_near_double_on_config1():
    """The float displacement path against the FP64 reference run, as the reference
    states its own tolerance (tests/fast_lattice_tests.cpp:435-447: max |du| < 1e-6 m)."""
    f, d = load("c1_float"), load("c1_double")
    du = np.abs(f["u"].astype(np.float64) - d["u"]).max()
    assert du < 1e-6
    assert np.array_equal(f["alive"], d["alive"])
