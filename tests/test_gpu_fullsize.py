"""The headline workloads at their full sizes on the device.

The config-2 impact (128x64x128: 1,048,576 nodes, 13,338,492 bonds) and the
config-5 plank cut (256x32x32: 262,144 nodes, 3,253,116 bonds) are checked
through what holds at any size:
- the first 14 substeps equal the oracle bit for bit (the striker's first
  contacts, the start sample of every node and the first substeps whose
  candidates overflow the block's list included; at this size the colour
  sweep runs as stage pairs and stage triples on LDS tiles);
- the first 4 substeps of the impact in double, and of a 96x48x128 block
  (whose stages group differently), equal the oracle bit for bit;
- a run repeats bit for bit;
- the device's fragment labels equal the host union-find on the same bonds;
- a batch slot equals the plain backend;
- the status counters agree with the downloaded state, and the fragments'
  mass properties add up to the body's mass and momentum.
"""
import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.runner import run_batch, run_scene
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu
F = abi.PRECISION_FLOAT


@pytest.fixture(scope="module")
def c2_run(gpu):
    sc = scenes.config2(substeps=120)
    print(f"config 2 N {sc.node_count} B {sc.bond_count}, 120 substeps", flush=True)
    return sc, run_scene(gpu, sc, F, labels="device", keep_backend=True)


def test_config2_first_substeps_equal_the_oracle(gpu, oracle):
    # The single block's candidate list first overflows its 1,024 entries in
    # substep 10; 14 substeps take the comparison four substeps into that regime.
    sc = scenes.config2(substeps=14)
    print("config 2, 14 substeps, HIP against the oracle", flush=True)
    h = run_scene(gpu, sc, F, labels="device")
    o = run_scene(oracle, sc, F)
    assert h["status"].contact.impulse_contacts > 0
    assert h["status"].contact.candidate_overflow > 0
    assert_same(h, o)


@pytest.mark.parametrize("dims,precision", [((128, 64, 128), abi.PRECISION_DOUBLE), ((96, 48, 128), F)],
                         ids=["config2-double", "96x48x128-float"])
def test_grouped_sweep_sizes_equal_the_oracle(gpu, oracle, dims, precision):
    # Lattices of at least 524,288 nodes sweep in stage pairs and, where every
    # tile's three-stage lists fit one workgroup, stage triples; which stages
    # group depends on the lattice's shape. The double lane's kernels hold
    # more registers per thread than the float lane's.
    sc = scenes.impact(*dims, substeps=4)
    print(f"impact {dims}, {'double' if precision == abi.PRECISION_DOUBLE else 'float'}, 4 substeps", flush=True)
    h = run_scene(gpu, sc, precision, labels="device")
    o = run_scene(oracle, sc, precision)
    assert h["status"].contact.impulse_contacts > 0
    assert_same(h, o)


def test_config2_repeats_and_labels_match_the_host(gpu, c2_run):
    sc, a = c2_run
    assert a["status"].broken_bonds > 0 and a["components"] >= 1
    b = run_scene(gpu, sc, F, labels="device")
    for k in ("u", "v", "alive", "damage", "mode", "labels"):
        assert np.array_equal(a[k], b[k]), k
    host, count = gpu.components(sc.node_count, sc.asset.bond_a, sc.asset.bond_b, a["alive"])
    assert count == a["components"] and np.array_equal(host, a["labels"])


def test_config2_counters_agree_with_the_state(gpu, c2_run):
    sc, a = c2_run
    st = a["status"]
    dead = a["alive"] == 0
    assert st.broken_bonds == int(dead.sum())
    assert np.all(dead[a["first_failure"]]) and len(a["first_failure"]) > 0
    assert np.all(a["mode"][dead] == 1)              # tensile only: compression and shear are off
    assert np.all(a["damage"][dead] >= 1.0) and st.max_damage >= 1.0
    assert st.removed_energy_j > 0.0 and st.contact.candidate_overflow > 0
    assert st.max_tensile_stretch > 0.0


def test_config2_fragment_mass_adds_up(gpu, c2_run):
    sc, a = c2_run
    be = a["backend"]
    try:
        frags = be.fragment_mass_properties(sc.cell)
    finally:
        be.close()
    mass = sc.asset.mass
    v = a["state"].v.reshape(-1, 3)
    assert len(frags) == a["components"]
    assert np.array_equal(frags["node_count"], np.bincount(a["labels"]))
    total = mass.sum()
    assert abs(frags["mass_kg"].sum() - total) <= 1e-12 * total
    p = (frags["mass_kg"][:, None] * frags["linear_velocity_m_s"]).sum(0)
    scale = (mass[:, None] * np.abs(v)).sum()
    assert np.all(np.abs(p - (mass[:, None] * v).sum(0)) <= 1e-9 * scale)


def test_config2_batch_slot_equals_plain(gpu, c2_run):
    sc, a = c2_run
    print("config 2 in a one-slot batch", flush=True)
    (b,) = run_batch(gpu, [sc], F)
    for k in ("u", "v", "alive", "damage", "labels"):
        assert np.array_equal(a[k], b[k]), k
    assert a["status"].broken_bonds == b["status"].broken_bonds


def test_config5_cut_labels_match_the_host(gpu):
    sc = scenes.config5(substeps=12 * 60)
    print(f"config 5 N {sc.node_count} B {sc.bond_count}, 720 substeps", flush=True)
    a = run_scene(gpu, sc, F, labels="device")
    assert a["status"].broken_bonds > 0
    host, count = gpu.components(sc.node_count, sc.asset.bond_a, sc.asset.bond_b, a["alive"])
    assert count == a["components"] and np.array_equal(host, a["labels"])
    b = run_scene(gpu, sc, F, labels="device")
    for k in ("u", "alive", "damage", "labels"):
        assert np.array_equal(a[k], b[k]), k
    assert a["status"].plastic_work_j == b["status"].plastic_work_j > 0.0
