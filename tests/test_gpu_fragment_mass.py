"""Fragment mass properties on the device (blx_backend_fragment_mass_properties,
blx_batch_fragment_mass_properties) against the oracle.

The device adds each fragment's terms in a fixed tree order. The reference
adds them serially in node order, and so does the oracle, which
tests/test_fragment_mass.py pins to the reference bit for bit. The two
therefore agree to rounding. Every field is held to TOL = 1e-9 of its natural
scale:
- the mass, for the mass;
- the largest node position, for the centre of mass;
- the largest node speed, for the linear velocity;
- the trace, for the inertia;
- the sum of m |r| |v - v_com| + m h^2/6 |spin|, for the angular momentum
  (and that over the smallest principal inertia for the angular velocity);
- the source kinetic energy, for the three energies.
Rounding differences of sums over 10^5 nodes stay below 1e-11 of these
scales. A dropped term, a wrong sign or a transposed index is off by order one.
The node count and first node of each fragment are exact. Runs repeat bit for
bit, and a batch slot gives the plain backend's bits.
"""
import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.lattice import Asset, Batch
from banjo_amd.runner import make_control, prepare, run_scene
from test_fragment_mass import lists

pytestmark = pytest.mark.gpu
F, D = abi.PRECISION_FLOAT, abi.PRECISION_DOUBLE
TOL = 1e-9


def world(state):
    """Node positions origin + x0 + u (FastLattice.cpp:179), velocities, masses."""
    pos = (state.origin[None, :] + state.x0.reshape(-1, 3)) + state.u.reshape(-1, 3)
    return pos, state.v.reshape(-1, 3).copy(), state.mass.copy()


def oracle_of(oracle, be, state, cell, spin=None):
    labels, count = be.fragment_labels()
    pos, vel, mass = world(state)
    off, nodes = lists(labels)
    return oracle.fragment_mass(pos, vel, mass, cell, off, nodes, spin), (pos, vel, mass, labels)


def assert_close(dev, orc, inputs, cell, spin=None):
    pos, vel, mass, labels = inputs
    assert len(dev) == len(orc)
    assert np.array_equal(dev["node_count"], orc["node_count"])
    assert np.array_equal(dev["first_node"], orc["first_node"])
    worst = 0.0
    for f in range(len(orc)):
        o, d = orc[f], dev[f]
        sel = labels == f
        p, v, m = pos[sel], vel[sel], mass[sel]
        r = p - o["center_of_mass_world_m"]
        cdi = m * cell * cell / 6.0
        sl = (m * np.linalg.norm(r, axis=1) * np.linalg.norm(v - o["linear_velocity_m_s"], axis=1)).sum()
        if spin is not None:
            sl += (cdi * np.linalg.norm(spin[sel], axis=1)).sum()
        smallest = np.linalg.eigvalsh(o["inertia_world_kg_m2"]).min()
        ke = max(o["source_kinetic_energy_j"], 1e-300)
        scales = {
            "mass_kg": o["mass_kg"],
            "center_of_mass_world_m": np.abs(p).max(),
            "linear_velocity_m_s": max(np.abs(v).max(), 1e-300),
            "inertia_world_kg_m2": np.trace(o["inertia_world_kg_m2"]),
            "angular_momentum_kg_m2_s": max(sl, 1e-300),
            "angular_velocity_rad_s": max(sl / smallest, 1e-300),
            "source_kinetic_energy_j": ke,
            "rigid_kinetic_energy_j": ke,
            "coarsening_kinetic_loss_j": ke,
        }
        for k, s in scales.items():
            err = np.max(np.abs(np.asarray(d[k], float) - np.asarray(o[k], float))) / s
            assert err <= TOL, (f, k, err, d[k], o[k])
            worst = max(worst, err)
    print(f"fragments {len(orc)} worst scaled difference {worst:.3e}", flush=True)


def run_kept(gpu, sc, precision, substeps=None):
    res = run_scene(gpu, sc, precision, substeps=substeps, labels=None, keep_backend=True)
    return res["backend"], res["state"]


CASES = {
    "cut_float": (lambda: scenes.plank_cut(24, 6, 6, speed=4.0, substeps=120, ledge_cells=4), F),
    "impact_4slabs_double": (lambda: scenes.small_impact(blocks=4, substeps=160), D),
    "bridge_float": (lambda: scenes.bridge_impact(), F),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_matches_oracle_after_a_run(gpu, oracle, name):
    sc, precision = CASES[name][0](), CASES[name][1]
    print(f"mass properties {name} N {sc.node_count} B {sc.bond_count}", flush=True)
    be, state = run_kept(gpu, sc, precision)
    try:
        dev = be.fragment_mass_properties(sc.cell)
        orc, inputs = oracle_of(oracle, be, state, sc.cell)
        assert len(orc) >= 1
        assert_close(dev, orc, inputs, sc.cell)
        again = be.fragment_mass_properties(sc.cell)
        assert again.tobytes() == dev.tobytes()
        # with a spin per node
        spin = np.random.default_rng(3).normal(0.0, 20.0, (sc.node_count, 3))
        dev = be.fragment_mass_properties(sc.cell, spin)
        orc, inputs = oracle_of(oracle, be, state, sc.cell, spin)
        assert_close(dev, orc, inputs, sc.cell, spin)
    finally:
        be.close()


def test_many_fragments_and_tiles(gpu, oracle):
    """A cut that shatters: hundreds of pieces, large ones crossing many tiles."""
    sc = scenes.small_impact(nx=40, ny=10, nz=40, ball_speed=40.0, substeps=200)
    print(f"mass properties shatter N {sc.node_count}", flush=True)
    be, state = run_kept(gpu, sc, F)
    try:
        dev = be.fragment_mass_properties(sc.cell)
        orc, inputs = oracle_of(oracle, be, state, sc.cell)
        print(f"fragments {len(orc)} largest {orc['node_count'].max()}", flush=True)
        assert len(orc) > 1 and orc["node_count"].max() > 4 * 256
        assert_close(dev, orc, inputs, sc.cell)
    finally:
        be.close()


def tiny(pos, vel, mass, bonds=()):
    pos = np.asarray(pos, float)
    n = len(pos)
    a = np.array([b[0] for b in bonds], np.uint32)
    b = np.array([b[1] for b in bonds], np.uint32)
    L = np.array([np.linalg.norm(pos[j] - pos[i]) for i, j in bonds]) if bonds else np.zeros(0)
    thr = np.tile([1.0, 2.0, 1.0, 2.0, 1.0, 2.0], (len(bonds), 1)) if bonds else np.zeros((0, 6))
    asset = Asset(reference_position=pos, mass=np.asarray(mass, float), bond_a=a, bond_b=b, rest_length=L,
                  compliance=np.full(len(bonds), 1e-6), threshold=thr, velocity=np.asarray(vel, float))
    return scenes.Scene(name="tiny", dims=(n, 1, 1), cell=0.01, asset=asset, slabs=np.array([0, n], np.uint32),
                        settings=scenes.settings(1e-5), sphere=scenes.sphere_state(center=(0.0, 100.0, 0.0)),
                        substeps=0)


def uploaded(gpu, sc, precision):
    sched, state = prepare(gpu, sc)
    be = gpu.backend(precision, sched, 0)
    be.upload(state, sc.settings, sc.sphere)
    return be


@pytest.mark.parametrize("precision", [F, D])
def test_known_answers_on_the_device(gpu, precision):
    # One spinning cell (tests/transfer_accounting_tests.cpp:59-73).
    spin = np.array([[2.0, -3.0, 4.0]])
    be = uploaded(gpu, tiny([[1.0, 2.0, 3.0]], [[0.2, 0.3, -0.4]], [0.02]), precision)
    f = be.fragment_mass_properties(0.002, spin)[0]
    be.close()
    inertia = 0.02 * 0.002 * 0.002 / 6.0
    assert np.linalg.norm(f["angular_momentum_kg_m2_s"] - inertia * spin[0]) <= 1e-20
    assert np.linalg.norm(f["angular_velocity_rad_s"] - spin[0]) <= 1e-12
    assert abs(f["coarsening_kinetic_loss_j"]) <= 1e-15
    # Two bonded cells moving apart (tests/transfer_accounting_tests.cpp:125-138).
    be = uploaded(gpu, tiny([[-0.2, 0, 0], [0.2, 0, 0]], [[-2.0, 0, 0], [2.0, 0, 0]], [1.0, 1.0], [(0, 1)]),
                  precision)
    f = be.fragment_mass_properties(0.01)
    be.close()
    assert len(f) == 1
    assert abs(f[0]["source_kinetic_energy_j"] - 4.0) <= 1e-12
    assert abs(f[0]["coarsening_kinetic_loss_j"] - 4.0) <= 1e-12


def test_rigid_motion_on_the_device(gpu):
    """A 40x20x20 box moving rigidly (v = boost + w x x, spin w): one fragment
    over 63 tiles recovers w and loses no energy to coarsening."""
    sc = scenes.cantilever(40, 20, 20, tip_load=0.0)
    w, boost = np.array([0.1, 0.2, -0.3]), np.array([0.03, -0.02, 0.01])
    pos = sc.asset.reference_position
    sc.asset.velocity = boost + np.cross(w, pos)
    sc.clamped = None
    be = uploaded(gpu, sc, D)
    try:
        f = be.fragment_mass_properties(sc.cell, np.tile(w, (sc.node_count, 1)))
    finally:
        be.close()
    assert len(f) == 1 and f[0]["node_count"] == sc.node_count
    assert np.linalg.norm(f[0]["angular_velocity_rad_s"] - w) <= 1e-10 * np.linalg.norm(w)
    assert abs(f[0]["coarsening_kinetic_loss_j"]) <= 1e-10 * f[0]["source_kinetic_energy_j"]


def test_batch_slot_gives_the_backends_bits(gpu):
    sc = scenes.plank_cut(20, 6, 5, speed=4.0, substeps=90, ledge_cells=4)
    be, _ = run_kept(gpu, sc, F)
    try:
        plain = be.fragment_mass_properties(sc.cell)
    finally:
        be.close()
    sched, state = prepare(gpu, sc)
    b = Batch(gpu, F, 2)
    try:
        for slot in (0, 1):
            b.upload(slot, sched, state, sc.settings, sc.sphere)
            b.arm(slot, make_control(sc.substeps))
        b.set_blade(sc.info["blade"])
        while b.advance(1 << 62):
            pass
        for slot in (0, 1):
            assert b.fragment_mass_properties(slot, sc.cell).tobytes() == plain.tobytes()
    finally:
        b.close()


def test_errors_on_the_device(gpu):
    be = uploaded(gpu, tiny([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0]], [[0, 0, 0], [0, 0, 0]], [1.0, 0.0]), D)
    try:
        with pytest.raises(RuntimeError, match="no mass"):
            be.fragment_mass_properties(0.01)  # the second node is a fragment of no mass
    finally:
        be.close()
    be = uploaded(gpu, tiny([[0.0, 0.0, 0.0]], [[0, 0, 0]], [1.0]), D)
    try:
        with pytest.raises(RuntimeError, match="singular"):
            be.fragment_mass_properties(0.0)  # a point mass of no extent
        with pytest.raises(ValueError):
            be.fragment_mass_properties(0.01, np.zeros(2))
    finally:
        be.close()
