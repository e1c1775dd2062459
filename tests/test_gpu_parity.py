"""HIP backend against the oracle restatement, bit for bit in float and double.

The oracle itself is pinned to the reference's own lattice lane (tests/test_oracle.py).
Every case runs the same seeded scene through both and requires identical
displacements, velocities, bond aliveness, damage, failure modes, fragment
labels and the exact counters of the run status. Ledgers the HIP backend sums
with a fixed-order tree (load ledger, plastic work) are compared at 1e-12
relative, the tolerance the reference's own GPU parity test uses for its
atomically summed energies (tests/fast_lattice_tests.cpp:520-522).
"""
import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.runner import run_scene

pytestmark = pytest.mark.gpu

EXACT_STATUS = ("total_steps", "first_failure_step", "last_failure_step", "failure_rounds", "broken_bonds",
                "removed_energy_j", "last_energy_gain_step", "max_damage", "last_damage_gain_step", "exit_reason",
                "max_tensile_stretch", "max_compressive_strain", "max_shear_strain", "rank_deficient_nodes")
EXACT_CONTACT = ("impulse_contacts", "impulse_to_material", "dissipated_kinetic_energy_j", "maximum_penetration_m",
                 "maximum_position_correction_m", "maximum_center_shift_m", "candidate_overflow", "ball_support_events")


def by_asset(res, name):
    """A per-bond state array in asset bond order (the state holds schedule order)."""
    arr = getattr(res["state"], name)
    out = np.empty_like(arr)
    out[res["schedule"].bond_order] = arr
    return out


def assert_same(a, b, ledger_rtol=1e-12):
    assert np.array_equal(a["u"], b["u"]), f"max |du| {np.abs(a['u'] - b['u']).max()}"
    assert np.array_equal(a["v"], b["v"]), f"max |dv| {np.abs(a['v'] - b['v']).max()}"
    assert np.array_equal(a["u_prev"], b["u_prev"])
    assert np.array_equal(a["alive"], b["alive"]), f"{(a['alive'] != b['alive']).sum()} alive mismatches"
    assert np.array_equal(a["damage"], b["damage"])
    assert np.array_equal(a["mode"], b["mode"])
    assert np.array_equal(a["labels"], b["labels"])
    assert np.array_equal(a["first_failure"], b["first_failure"])
    if "state" in a and "state" in b:
        for name in ("prev_tensile", "prev_compressive", "prev_shear", "plastic_extension", "plastic_strain"):
            assert np.array_equal(by_asset(a, name), by_asset(b, name)), name
    sa, sb = a["status"], b["status"]
    for f in EXACT_STATUS:
        assert getattr(sa, f) == getattr(sb, f), f
    for f in EXACT_CONTACT:
        va, vb = getattr(sa.contact, f), getattr(sb.contact, f)
        assert (list(va) if hasattr(va, "__len__") else va) == (list(vb) if hasattr(vb, "__len__") else vb), f
    la, lb = sa.external_load, sb.external_load
    assert la.steps == lb.steps
    for f in ("requested_impulse_n_s", "impulse_n_s", "requested_angular_impulse_kg_m2_s", "angular_impulse_kg_m2_s"):
        np.testing.assert_allclose(list(getattr(la, f)), list(getattr(lb, f)), rtol=ledger_rtol, atol=1e-300)
    np.testing.assert_allclose(la.work_j, lb.work_j, rtol=ledger_rtol)
    np.testing.assert_allclose(sa.plastic_work_j, sb.plastic_work_j, rtol=ledger_rtol)
    for k in range(3):
        assert getattr(a["sphere"], "center")[k] == getattr(b["sphere"], "center")[k]
        assert getattr(a["sphere"], "velocity")[k] == getattr(b["sphere"], "velocity")[k]


def _iterations(sc, n):
    sc.settings.constraint_iterations = n
    return sc


CASES = {
    "cantilever16_300": lambda: scenes.config1(substeps=300),
    "impact_1block": lambda: scenes.small_impact(substeps=150),
    "impact_4slabs": lambda: scenes.small_impact(substeps=150, blocks=4),
    # More striker candidates than a block lists: the capped, ordered lists and
    # the overflow count of CpuLatticeBackend.cpp:377-385 on every contact substep.
    "overflow_1block": lambda: scenes.overflow_impact(),
    "overflow_4slabs": lambda: scenes.overflow_impact(blocks=4),
    # Two constraint iterations: the carried multiplier and the projection between them.
    "impact_2iterations": lambda: scenes.overflow_impact(substeps=20, iterations=2),
    "bridge_2iterations": lambda: _iterations(scenes.bridge_impact(substeps=80), 2),
}


@pytest.mark.parametrize("precision", [abi.PRECISION_FLOAT, abi.PRECISION_DOUBLE], ids=["float", "double"])
@pytest.mark.parametrize("case", list(CASES))
def test_hip_matches_oracle(gpu, oracle, case, precision):
    sc = CASES[case]()
    print(f"case {case} precision {precision} N {sc.node_count} B {sc.bond_count}", flush=True)
    o = run_scene(oracle, sc, precision)
    # These lattices are small enough to run a striker substep on one stream;
    # streams 1 also runs them split across streams, as a large lattice runs.
    for tuning in ({}, {"streams": 1}):
        print(f"  tuning {tuning}", flush=True)
        h = run_scene(gpu, sc, precision, labels="device", tuning=tuning)
        assert h["status"].broken_bonds > 0
        if case.startswith("overflow"):
            assert h["status"].contact.candidate_overflow > 1000
        assert_same(h, o)
