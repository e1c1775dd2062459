"""Fragment mass properties: the oracle against the reference and against what
the mathematics fixes.

calculateFragmentMassProperties (src/fracture/FragmentMassProperties.cpp:7-80)
is restated in the oracle (orc_fragment_mass). Here it is pinned three ways:

* bit for bit against the reference's own function compiled into oracle/_ref
  (ref_fragment_mass), on random fragments with and without spin;
* against the known answers of the reference's tests: one spinning cell keeps
  its spin and carries m h^2/6 of inertia (tests/transfer_accounting_tests.cpp:59-73),
  two cells moving apart lose all their kinetic energy to coarsening
  (:125-138), and a rigidly moving body keeps its angular velocity and loses
  nothing (tests/energy_rupture_tests.cpp:64-76);
* against conservation over any partition: the fragments' masses, momenta
  and angular momenta about the origin add up to the whole body's
  (tests/energy_rupture_tests.cpp:47,76).

Each known answer fails under a plausible slip: the spinning cell needs the
finite-cell term in both the inertia and the angular momentum; the rigid body
needs the signs of the off-diagonal inertia and the orientation of the cross
product; the two cells need the centre-of-mass velocity subtracted.
"""
import numpy as np
import pytest


def lists(labels):
    """Node lists of a labelling, as findConnectedComponents hands them over:
    ascending nodes per fragment, fragments in label order."""
    order = np.argsort(labels, kind="stable").astype(np.uint32)
    counts = np.bincount(labels, minlength=int(labels.max()) + 1)
    return np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32), order


def random_body(seed, n, spin=True):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.3, 0.3, (n, 3)) + np.array([0.7, -0.2, 1.1])
    vel = rng.normal(0.0, 2.0, (n, 3))
    mass = rng.uniform(0.5e-3, 2e-3, n)
    sp = rng.normal(0.0, 5.0, (n, 3)) if spin else None
    return pos, vel, mass, sp


def random_labels(seed, n, k):
    rng = np.random.default_rng(seed + 1)
    lab = rng.integers(0, k, n)
    lab[: min(n, k)] = np.arange(min(n, k))  # every label used
    return lab.astype(np.uint32)


def whole_body(pos, vel, mass, sp, cell):
    """Mass, momentum and angular momentum about the origin of the nodes
    (finite cells spin with m h^2 / 6)."""
    m = mass.sum()
    p = (mass[:, None] * vel).sum(0)
    h = np.cross(pos, mass[:, None] * vel).sum(0)
    if sp is not None:
        h = h + ((mass * cell * cell / 6.0)[:, None] * sp).sum(0)
    return m, p, h


@pytest.mark.parametrize("spin", [True, False])
@pytest.mark.parametrize("n,k", [(1, 1), (7, 3), (500, 1), (2000, 37), (3000, 3000)])
def test_oracle_matches_reference_bit_for_bit(oracle, ref, n, k, spin):
    pos, vel, mass, sp = random_body(n * 31 + k, n, spin)
    off, nodes = lists(random_labels(n + k, n, k))
    cell = 0.01
    a = oracle.fragment_mass(pos, vel, mass, cell, off, nodes, sp)
    b = ref.fragment_mass(pos, vel, mass, cell, off, nodes, sp)
    assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("lib_name", ["oracle", "ref"])
def test_known_answers(request, lib_name):
    lib = request.getfixturevalue(lib_name)
    # One spinning cell (transfer_accounting_tests.cpp:59-73).
    spin = np.array([2.0, -3.0, 4.0])
    f = lib.fragment_mass([[1.0, 2.0, 3.0]], [[0.2, 0.3, -0.4]], [0.02], 0.002, [0, 1], [0], [spin])[0]
    inertia = 0.02 * 0.002 * 0.002 / 6.0
    assert np.linalg.norm(f["angular_momentum_kg_m2_s"] - inertia * spin) <= 1e-20
    assert np.linalg.norm(f["angular_velocity_rad_s"] - spin) <= 1e-12
    assert abs(f["coarsening_kinetic_loss_j"]) <= 1e-15
    assert np.allclose(f["inertia_world_kg_m2"], inertia * np.eye(3), rtol=1e-12, atol=0)
    # Two cells moving apart (transfer_accounting_tests.cpp:125-138).
    f = lib.fragment_mass([[-0.2, 0, 0], [0.2, 0, 0]], [[-2.0, 0, 0], [2.0, 0, 0]], [1.0, 1.0], 0.01,
                          [0, 2], [0, 1])[0]
    assert abs(f["source_kinetic_energy_j"] - 4.0) <= 1e-12
    assert abs(f["rigid_kinetic_energy_j"]) <= 1e-12
    assert abs(f["coarsening_kinetic_loss_j"] - 4.0) <= 1e-12
    assert f["node_count"] == 2 and f["first_node"] == 0


@pytest.mark.parametrize("lib_name", ["oracle", "ref"])
def test_rigid_motion_keeps_spin_and_energy(request, lib_name):
    """v = boost + w x x and spin w for every cell: the fragment recovers w and
    its rigid kinetic energy is all of the source energy."""
    lib = request.getfixturevalue(lib_name)
    g = np.stack(np.meshgrid(np.arange(6), np.arange(4), np.arange(5), indexing="ij"), -1).reshape(-1, 3)
    pos = 0.01 * g + np.array([0.3, -0.1, 0.2])
    w, boost = np.array([0.1, 0.2, -0.3]), np.array([0.03, -0.02, 0.01])
    vel = boost + np.cross(w, pos)
    n = len(pos)
    f = lib.fragment_mass(pos, vel, np.full(n, 1e-3), 0.01, [0, n], np.arange(n), np.tile(w, (n, 1)))[0]
    assert np.linalg.norm(f["angular_velocity_rad_s"] - w) <= 1e-12 * np.linalg.norm(w)
    assert np.linalg.norm(f["linear_velocity_m_s"] - (boost + np.cross(w, f["center_of_mass_world_m"]))) <= 1e-14
    assert abs(f["coarsening_kinetic_loss_j"]) <= 1e-12 * f["source_kinetic_energy_j"]


@pytest.mark.parametrize("k", [1, 5, 60])
def test_fragments_conserve_the_body(oracle, k):
    pos, vel, mass, sp = random_body(11 + k, 900)
    cell = 0.01
    off, nodes = lists(random_labels(5 * k, 900, k))
    frags = oracle.fragment_mass(pos, vel, mass, cell, off, nodes, sp)
    m, p, h = whole_body(pos, vel, mass, sp, cell)
    fm = frags["mass_kg"]
    fp = fm[:, None] * frags["linear_velocity_m_s"]
    fh = frags["angular_momentum_kg_m2_s"] + np.cross(frags["center_of_mass_world_m"], fp)
    assert abs(fm.sum() - m) <= 1e-12 * m
    assert np.linalg.norm(fp.sum(0) - p) <= 1e-12 * np.abs(mass[:, None] * vel).sum()
    assert np.linalg.norm(fh.sum(0) - h) <= 1e-11 * np.abs(np.cross(pos, mass[:, None] * vel)).sum()
    assert np.all(frags["rigid_kinetic_energy_j"] <= frags["source_kinetic_energy_j"] * (1 + 1e-12))
    assert np.array_equal(frags["node_count"], np.diff(off))


@pytest.mark.parametrize("lib_name", ["oracle", "ref"])
def test_errors(request, lib_name):
    lib = request.getfixturevalue(lib_name)
    with pytest.raises(RuntimeError, match="no mass"):
        lib.fragment_mass([[0, 0, 0]], [[0, 0, 0]], [0.0], 0.01, [0, 1], [0])
    with pytest.raises(RuntimeError, match="singular"):
        lib.fragment_mass([[0, 0, 0]], [[0, 0, 0]], [1.0], 0.0, [0, 1], [0])  # a point mass of no extent
    with pytest.raises(ValueError):
        lib.fragment_mass([[0, 0, 0]], [[0, 0, 0]], [1.0], 0.01, [0, 0], [])  # an empty fragment
