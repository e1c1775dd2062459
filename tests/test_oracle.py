"""Pins the oracle restatement.

1. Against the reference's own lattice lane (oracle/_ref, compiled from the
   reference sources): schedule, state, every substep of seeded runs through
   fracture, in float and double, on one block and on four slabs, bit for bit.
2. Against what the mathematics fixes, independent of the reference: a single
   XPBD spring is backward Euler (tests/lattice_verlet_tests.cpp:45-55 pins the
   same law), the plastic return's closed form
   (tests/lattice_plasticity_tests.cpp:191-233), an affine deformation's
   Green-Lagrange strain and a rigid rotation's zero strain
   (tests/fast_lattice_tests.cpp:165-224), the step damage law, and
   brute-force fragment labels with the canonical order
   (fracture/ConnectedComponents.cpp:77-85).
Each check is chosen so that a plausible slip in the restatement (a dropped
term, a sign, a transposed product, a wrong sweep order) fails it.
"""
import math

import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.runner import prepare, run_scene
from lattice_cases import affine_box, green_lagrange, scene_of, two_node

F, D = abi.PRECISION_FLOAT, abi.PRECISION_DOUBLE


def same_run(a, b, first_failure=True, skip=()):
    for k in ("u", "v", "u_prev", "alive", "damage", "mode") + (("first_failure",) if first_failure else ()):
        assert np.array_equal(a[k], b[k]), k
    for f in ("total_steps", "broken_bonds", "failure_rounds", "first_failure_step", "last_failure_step",
              "removed_energy_j", "max_damage", "last_damage_gain_step", "last_energy_gain_step", "exit_reason",
              "max_tensile_stretch", "max_compressive_strain", "max_shear_strain", "rank_deficient_nodes",
              "plastic_work_j"):
        if f not in skip:
            assert getattr(a["status"], f) == getattr(b["status"], f), f
    for f in ("impulse_contacts", "dissipated_kinetic_energy_j", "maximum_penetration_m", "candidate_overflow",
              "maximum_position_correction_m", "maximum_center_shift_m", "ball_support_events"):
        assert getattr(a["status"].contact, f) == getattr(b["status"].contact, f), f
    assert list(a["sphere"].center) == list(b["sphere"].center)


# ---------------------------------------------------------------------------
# 1. Against the reference build.

def test_schedule_and_state_match_reference(oracle, ref):
    for sc in (scenes.small_impact(blocks=4), scenes.cantilever(9, 7, 11, blocks=3)):
        so, sr = oracle.schedule_build(sc.asset, sc.slabs), ref.schedule_build(sc.asset, sc.slabs)
        for f in ("bond_order", "bond_schedule_index", "bond_block_begin", "range_begin", "range_end", "bond_color"):
            assert np.array_equal(getattr(so, f), getattr(sr, f)), f
        assert so.color_count == sr.color_count == 26
        to, tr = oracle.state_build(sc.asset, sr, (0.5, 1.0, -2.0)), ref.state_build(sc.asset, sr, (0.5, 1.0, -2.0))
        for k, v in tr.__dict__.items():
            if isinstance(v, np.ndarray):
                assert np.array_equal(v, getattr(to, k)), k


RUNS = {
    "cantilever16_float": (lambda: scenes.config1(substeps=120), F, False),
    "cantilever16_double_direct": (lambda: scenes.config1(substeps=120), D, True),
    "impact_1block_float": (lambda: scenes.small_impact(nx=16, ny=10, nz=16, substeps=100), F, False),
    "impact_4slabs_float": (lambda: scenes.small_impact(nx=16, ny=10, nz=16, substeps=100, blocks=4), F, False),
    "impact_4slabs_double": (lambda: scenes.small_impact(nx=16, ny=10, nz=16, substeps=100, blocks=4), D, False),
    "plastic_float": (lambda: scenes.plastic_impact(substeps=80), F, False),
    "plastic_double_4slabs": (lambda: scenes.plastic_impact(blocks=4, substeps=60), D, False),
    "bridge_ledges_float": (lambda: scenes.bridge_impact(substeps=100), F, False),
    "bridge_ledges_3slabs_double": (lambda: scenes.bridge_impact(blocks=3, substeps=80), D, True),
    # More striker candidates than a block lists (1,024): the overflow branch
    # of CpuLatticeBackend.cpp:377-385 on every substep of the contact.
    "overflow_1block_float": (lambda: scenes.overflow_impact(), F, False),
    "overflow_1block_double": (lambda: scenes.overflow_impact(), D, False),
    "overflow_4slabs_float": (lambda: scenes.overflow_impact(blocks=4), F, False),
    "overflow_4slabs_double": (lambda: scenes.overflow_impact(blocks=4), D, False),
    # Two constraint iterations: the multiplier carried between them
    # (LatticePhysics.hpp:1160-1163) and the support projection after each.
    "impact_2iterations_float": (lambda: scenes.overflow_impact(substeps=20, iterations=2), F, False),
    "bridge_2iterations_double": (lambda: _iterations(scenes.bridge_impact(substeps=80), 2), D, False),
}


def _iterations(sc, n):
    sc.settings.constraint_iterations = n
    return sc


@pytest.mark.parametrize("name", list(RUNS))
def test_runs_match_reference_bit_for_bit(oracle, ref, name):
    make, precision, direct = RUNS[name]
    sc = make()
    sc.settings.direct_arithmetic = 1 if direct else 0
    o = run_scene(oracle, sc, precision)
    r = run_scene(ref, sc, precision, backend_args=(0, 1))
    assert r["status"].broken_bonds > 0
    if name.startswith("overflow"):
        assert r["status"].contact.candidate_overflow > 1000
    same_run(o, r)
    assert np.array_equal(o["labels"], r["labels"])


def test_parallel_reference_equals_serial(ref):
    sc = scenes.small_impact(nx=16, ny=10, nz=16, substeps=60, blocks=4)
    a = run_scene(ref, sc, F, backend_args=(0, 1))
    b = run_scene(ref, sc, F, backend_args=(1, 4))
    # The threaded reference records no first-failure list, books last_energy_gain_step
    # per round (ParallelCpuLatticeBackend.cpp:1914-1919) and counts rank-deficient
    # recomputations per substep; every array and the other counters agree.
    same_run(a, b, first_failure=False, skip=("last_energy_gain_step", "rank_deficient_nodes"))


def test_runs_continue_across_calls(oracle, ref):
    sc = scenes.small_impact(nx=14, ny=8, nz=14, substeps=90)
    whole = run_scene(oracle, sc, F)
    parts = run_scene(oracle, sc, F, chunks=[7, 40, 43])
    assert np.array_equal(whole["u"], parts["u"]) and np.array_equal(whole["alive"], parts["alive"])
    same_run(parts, run_scene(ref, sc, F, backend_args=(0, 1), chunks=[7, 40, 43]))


@pytest.mark.parametrize("kw,reason", [
    (dict(calm_steps=5), 5), (dict(no_failure_steps=4), 2), (dict(quiet_steps=3, min_steps=2), 1),
    (dict(energy_flat_steps=2, min_steps=1), 4), (dict(removable_energy_j=1e-3), 6), ({}, 3)])
def test_exit_reasons_match_reference(oracle, ref, kw, reason):
    sc = scenes.small_impact(nx=12, ny=8, nz=12, substeps=80) if reason in (1, 4, 6, 3) else scenes.config1(substeps=20)
    if reason in (2, 5):
        sc.forces = None  # nothing loads the bar: calm, nothing fails
    o = run_scene(oracle, sc, F, control_kw=kw)
    r = run_scene(ref, sc, F, backend_args=(0, 1), control_kw=kw)
    assert o["status"].exit_reason == reason
    same_run(o, r)


def test_frames_match_reference(oracle, ref):
    sc = scenes.small_impact(nx=10, ny=6, nz=10, substeps=25)
    out = []
    for lib, extra in ((oracle, ()), (ref, (0, 1))):
        res = run_scene(lib, sc, F, backend_args=extra, control_kw=dict(capture_stride=7, max_frames=3),
                        keep_backend=True)
        out.append(res["backend"].frames())
        res["backend"].close()
    assert [f["step"] for f in out[0]] == [0, 7, 14] == [f["step"] for f in out[1]]
    for a, b in zip(*out):
        assert np.array_equal(a["u"], b["u"]) and np.array_equal(a["alive"], b["alive"])
        assert np.array_equal(a["damage"], b["damage"])


def test_refusals_match_reference(oracle, ref):
    sc = scenes.config1(substeps=1)
    for lib, extra in ((oracle, ()), (ref, (0, 1))):
        sched, state = prepare(lib, sc)
        be = lib.backend(F, sched, *extra)
        be.upload(state, sc.settings, sc.sphere)
        bad = np.zeros((sc.node_count, 3))
        bad[0, 1] = -1.0  # node 0 sits on the clamped face: inv_mass 0
        with pytest.raises(ValueError):
            be.set_external_forces(bad, 10)
        bad[0, 1] = math.inf
        with pytest.raises(ValueError):
            be.set_external_forces(bad, 10)
        with pytest.raises(ValueError):
            be.set_external_forces(np.zeros((sc.node_count, 3)) + 1.0, 0)


# ---------------------------------------------------------------------------
# 2. Against the mathematics.

@pytest.mark.parametrize("direct", [False, True])
def test_single_spring_is_backward_euler(oracle, direct):
    """XPBD with one iteration on a spring with one end fixed: q' = (q + dt v)/(1 + dt^2 k/m)."""
    L0, c, m, dt = 0.01, 1e-6, 1e-3, 1e-4
    asset = two_node(L0, c, m)
    q, v = 2e-5, 0.3
    pos = asset.reference_position.copy()
    pos[1, 0] += q
    vel = np.zeros((2, 3))
    vel[1, 0] = v
    sc = scene_of(asset, dt, clamped=np.array([True, False]), velocity=vel, position=pos, direct=direct)
    r = run_scene(oracle, sc, D)
    q_new = r["u"][1, 0]
    expected = (q + dt * v) / (1.0 + dt * dt / (c * m))
    assert q_new == pytest.approx(expected, rel=1e-12, abs=1e-20)
    assert r["v"][1, 0] == pytest.approx((expected - q) / dt, rel=1e-9)
    assert r["u"][0, 0] == 0.0


def test_plastic_return_closed_form(oracle):
    """Both ends fixed, bond held past yield: p = (e - yL)/(1 + H), work (yL + (yL + H p))/2 * p / c."""
    L0, c, y, H = 0.01, 1e-6, 0.004, 0.1
    asset = two_node(L0, c)
    e = 0.012 * L0
    pos = asset.reference_position.copy()
    pos[1, 0] += e
    sc = scene_of(asset, 1e-4, clamped=np.array([True, True]), position=pos, yield_stretch=y, hardening=H)
    r = run_scene(oracle, sc, D)
    p = (e - y * L0) / (1.0 + H)
    assert r["state"].plastic_extension[0] == pytest.approx(p, rel=1e-9)
    assert r["state"].plastic_strain[0] == pytest.approx(p, rel=1e-9)
    work = 0.5 * (y * L0 + (y * L0 + H * p)) * p / c
    assert r["status"].plastic_work_j == pytest.approx(work, rel=1e-9)
    # A second substep at the same length is inside the hardened yield surface: no further flow.
    sc.substeps = 2
    r2 = run_scene(oracle, sc, D)
    assert r2["state"].plastic_extension[0] == pytest.approx(p, rel=1e-9)


def _expected_peaks(asset, F):
    """Per bond: max(own stretch, d.E.d) with E the exact Green-Lagrange strain of F."""
    E = green_lagrange(F)
    ref = asset.reference_position
    d = ref[asset.bond_b] - ref[asset.bond_a]
    n = d / np.linalg.norm(d, axis=1)[:, None]
    stretch = np.linalg.norm(d @ np.asarray(F).T, axis=1) / asset.rest_length - 1.0
    normal = np.einsum("bi,ij,bj->b", n, E, n)
    return np.maximum(stretch, normal)


def test_affine_stretch_gives_green_lagrange_strain(oracle):
    e = 0.004
    Fm = np.diag([1.0 + e, 1.0 - 0.3 * e, 1.0 - 0.3 * e])
    Fm[0, 1] = 0.2 * e
    asset = affine_box(5, F=Fm)
    sc = scene_of(asset, 1e-9, direct=True)
    r = run_scene(oracle, sc, D)
    got = r["state"].prev_tensile[np.argsort(r["schedule"].bond_order)]
    want = _expected_peaks(asset, Fm)
    np.testing.assert_allclose(got, np.maximum(want, 0.0), rtol=1e-7, atol=1e-12)


def test_rigid_rotation_is_not_a_strain(oracle):
    th = 0.7
    Rm = np.array([[math.cos(th), -math.sin(th), 0.0], [math.sin(th), math.cos(th), 0.0], [0.0, 0.0, 1.0]])
    asset = affine_box(4, F=Rm, thresholds=np.full((0, 6), 0.0))
    B = asset.bond_count
    asset.threshold = np.tile([1e-6, 1e-6, 1e-6, 1e-6, 1e-6, 1e-6], (B, 1))
    for precision, direct, tol in ((D, True, 1e-12), (D, False, 1e-12), (F, False, 2e-6)):
        sc = scene_of(affine_box(4, F=Rm, thresholds=asset.threshold), 1e-9, direct=direct)
        r = run_scene(oracle, sc, precision)
        assert r["status"].broken_bonds == 0
        assert max(r["state"].prev_tensile.max(), r["state"].prev_shear.max(), r["state"].prev_compressive.max()) < tol


def test_step_damage_breaks_exactly_the_bonds_past_threshold(oracle):
    e = 0.005
    Fm = np.diag([1.0 + e, 1.0, 1.0])
    asset = affine_box(5, F=Fm)
    peaks = _expected_peaks(asset, Fm)
    thr = np.full((asset.bond_count, 6), math.inf)
    cut = 0.5 * (peaks.max() + np.unique(np.round(peaks, 9))[-2])
    thr[:, 0] = thr[:, 1] = cut
    asset.threshold = thr
    r = run_scene(oracle, scene_of(asset, 1e-9, direct=True), D)
    broke = r["alive"] == 0
    assert broke.sum() > 0
    assert np.array_equal(broke, peaks > cut)
    assert set(np.unique(r["mode"][broke])) == {1}


def test_damage_ramp_is_linear_between_start_and_end(oracle):
    e = 0.004
    Fm = np.diag([1.0 + e, 1.0, 1.0])
    asset = affine_box(4, F=Fm)
    peaks = np.maximum(_expected_peaks(asset, Fm), 0.0)
    thr = np.full((asset.bond_count, 6), math.inf)
    thr[:, 0] = 0.0
    thr[:, 1] = 0.02
    asset.threshold = thr
    r = run_scene(oracle, scene_of(asset, 1e-9, direct=True), D)
    np.testing.assert_allclose(r["damage"], np.clip(peaks / 0.02, 0, 1), rtol=1e-6, atol=1e-12)


def _bfs_labels(n, a, b, alive):
    adj = [[] for _ in range(n)]
    for x, y, al in zip(a, b, alive):
        if al:
            adj[x].append(y)
            adj[y].append(x)
    comp = [-1] * n
    comps = []
    for s in range(n):
        if comp[s] >= 0:
            continue
        stack, members = [s], []
        comp[s] = len(comps)
        while stack:
            x = stack.pop()
            members.append(x)
            for y in adj[x]:
                if comp[y] < 0:
                    comp[y] = comp[s]
                    stack.append(y)
        comps.append(members)
    order = sorted(range(len(comps)), key=lambda k: (-len(comps[k]), min(comps[k])))
    rank = {k: i for i, k in enumerate(order)}
    return np.array([rank[comp[i]] for i in range(n)], np.uint32), len(comps)


@pytest.mark.parametrize("seed", range(5))
def test_fragment_labels_brute_force(oracle, hip, seed):
    rng = np.random.default_rng(seed)
    n = 60
    m = 90
    a = rng.integers(0, n, m).astype(np.uint32)
    b = rng.integers(0, n, m).astype(np.uint32)
    keep = a != b
    a, b = a[keep], b[keep]
    alive = (rng.random(a.shape[0]) < 0.6).astype(np.uint8)
    want, count = _bfs_labels(n, a, b, alive)
    for lib in (oracle, hip):
        got, c = lib.components(n, a, b, alive)
        assert c == count and np.array_equal(got, want), lib.name
