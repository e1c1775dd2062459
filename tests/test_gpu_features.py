"""HIP backend feature coverage against the oracle and the golden fixtures.

Each case runs the same seeded scene through the HIP library and the oracle
(or reads the fixture written from the reference build) and requires the
same bits, as tests/test_gpu_parity.py does, across the rest of the lane's
surface: the full config-1 run, plasticity, radial damping, two support
planes with finite reach-capped footprints, direct arithmetic in float,
multi-call continuity, frames, exit rules, batch sizes, repeatability, the
load ledger, and the refusals the reference's GPU lane makes.
"""
import math

import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.lattice import control
from banjo_amd.runner import prepare, run_scene
from test_gpu_parity import assert_same
from test_golden import check_against_golden, run_case

pytestmark = pytest.mark.gpu
F, D = abi.PRECISION_FLOAT, abi.PRECISION_DOUBLE


@pytest.mark.parametrize("name", ["c1_float", "c1_double", "impact_4slabs"])
def test_hip_reproduces_golden(gpu, name):
    print(f"golden {name}", flush=True)
    res, g = run_case(gpu, name, labels="device")
    check_against_golden(res, g)


FEATURES = {
    "plastic_float": (lambda: scenes.plastic_impact(), F),
    "plastic_double_4slabs": (lambda: scenes.plastic_impact(blocks=4, substeps=80), D),
    "bridge_ledges_float": (lambda: scenes.bridge_impact(), F),
    "bridge_ledges_3slabs_double": (lambda: scenes.bridge_impact(blocks=3, substeps=100), D),
    # plank on two ledges, plastic, cut by a kinematic blade (blx_blade)
    "blade_cut_float": (lambda: scenes.plank_cut(24, 6, 6, speed=4.0, substeps=120, ledge_cells=4), F),
    "blade_cut_double": (lambda: scenes.plank_cut(20, 6, 5, speed=4.0, substeps=90, ledge_cells=4), D),
}


@pytest.mark.parametrize("name", list(FEATURES))
def test_feature_parity(gpu, oracle, name):
    sc, precision = FEATURES[name][0](), FEATURES[name][1]
    print(f"feature {name} N {sc.node_count} B {sc.bond_count}", flush=True)
    h = run_scene(gpu, sc, precision, labels="device")
    o = run_scene(oracle, sc, precision)
    assert h["status"].broken_bonds > 0
    assert_same(h, o)


def test_damping_and_direct_float(gpu, oracle):
    sc = scenes.small_impact(nx=14, ny=8, nz=14, substeps=90)
    sc.settings.damping_fraction = 1.0 - math.exp(-40.0 * sc.settings.dt)
    sc.settings.direct_arithmetic = 1
    h = run_scene(gpu, sc, F, labels="device")
    o = run_scene(oracle, sc, F)
    assert_same(h, o)


def test_plastic_flow_happens(gpu):
    sc = scenes.plastic_impact(substeps=60)
    h = run_scene(gpu, sc, F)
    assert h["status"].plastic_work_j > 0.0 and h["status"].max_plastic_stretch > 0.0
    assert np.abs(h["state"].plastic_extension).max() > 0.0


def test_runs_continue_across_calls_and_batches(gpu, oracle):
    sc = scenes.small_impact(nx=14, ny=8, nz=14, substeps=90)
    whole = run_scene(oracle, sc, F)
    for chunks, spl in (([7, 40, 43], 0), ([90], 37), ([1] * 5 + [85], 3)):
        h = run_scene(gpu, sc, F, chunks=chunks, control_kw=dict(steps_per_launch=spl), labels="device")
        assert_same(h, whole)
    h = run_scene(gpu, sc, F, control_kw=dict(steps_per_launch=37))
    assert h["status"].launches == (90 + 36) // 37


def test_repeatable_bit_for_bit(gpu):
    sc = scenes.small_impact(substeps=120)
    a = run_scene(gpu, sc, F, labels="device")
    b = run_scene(gpu, sc, F, labels="device")
    for k in ("u", "v", "alive", "damage", "labels"):
        assert np.array_equal(a[k], b[k]), k
    assert a["status"].removed_energy_j == b["status"].removed_energy_j


@pytest.mark.parametrize("kw,reason", [
    (dict(calm_steps=5), 5), (dict(no_failure_steps=4), 2), (dict(quiet_steps=3, min_steps=2), 1),
    (dict(energy_flat_steps=2, min_steps=1), 4), (dict(removable_energy_j=1e-3), 6), ({}, 3)])
def test_exit_reasons(gpu, oracle, kw, reason):
    sc = scenes.small_impact(nx=12, ny=8, nz=12, substeps=80) if reason in (1, 4, 6, 3) else scenes.config1(substeps=20)
    if reason in (2, 5):
        sc.forces = None
    h = run_scene(gpu, sc, F, control_kw=kw, labels="device")
    o = run_scene(oracle, sc, F, control_kw=kw)
    assert h["status"].exit_reason == reason
    assert_same(h, o)


def test_abandon_stops_between_batches(gpu):
    import ctypes
    sc = scenes.small_impact(nx=10, ny=6, nz=10, substeps=100)
    flag = (ctypes.c_uint8 * 1)(1)
    h = run_scene(gpu, sc, F, control_kw=dict(steps_per_launch=10, abandon=ctypes.cast(flag, ctypes.POINTER(ctypes.c_uint8))))
    assert h["status"].exit_reason == 7
    assert h["status"].total_steps == 10


def test_frames(gpu, oracle):
    sc = scenes.small_impact(nx=10, ny=6, nz=10, substeps=25)
    out = []
    for lib in (gpu, oracle):
        res = run_scene(lib, sc, F, control_kw=dict(capture_stride=7, max_frames=3), keep_backend=True)
        out.append(res["backend"].frames())
        res["backend"].close()
    assert [f["step"] for f in out[0]] == [0, 7, 14] == [f["step"] for f in out[1]]
    for a, b in zip(*out):
        assert np.array_equal(a["u"], b["u"]) and np.array_equal(a["alive"], b["alive"])
        assert np.array_equal(a["damage"], b["damage"])


def test_load_ledger_and_span(gpu, oracle):
    sc = scenes.config1(substeps=60)
    sc.force_substeps = 25  # the load expires part-way through the run
    h = run_scene(gpu, sc, F, labels="device")
    o = run_scene(oracle, sc, F)
    assert h["status"].external_load.steps == 25
    assert_same(h, o)


def test_refusals(gpu):
    sc = scenes.config1(substeps=1)
    sched, state = prepare(gpu, sc)
    be = gpu.backend(F, sched, 0)
    for field, value in (("bond_integrator", 1), ("plate_half_thickness", 0.005)):
        s = scenes.settings(sc.settings.dt)
        setattr(s, field, value)
        with pytest.raises(ValueError):
            be.upload(state, s, sc.sphere)
    for mode, mask in ((3, 63), (2, 5)):  # an unknown contact mode; a bucket mask that is no power of two minus one
        s = scenes.settings(sc.settings.dt)
        s.node_contact.mode, s.node_contact.bucket_mask = mode, mask
        with pytest.raises(ValueError):
            be.upload(state, s, sc.sphere)
    with pytest.raises(AssertionError):
        be.run(control(1))  # run before a successful upload: std::logic_error
    be.upload(state, sc.settings, sc.sphere)
    bad = np.zeros((sc.node_count, 3))
    bad[0, 1] = -1.0  # a clamped node
    with pytest.raises(ValueError):
        be.set_external_forces(bad, 10)
    be.close()


def test_phase_seconds_on_every_run(gpu):
    # RunStatus::phase_seconds (FastLattice.hpp:225-228) without profiling: each
    # device batch times one substep phase by phase and splits its device time
    # in those proportions, so the phases sum to the run's kernel time.
    sc = scenes.small_impact(nx=12, ny=8, nz=12, substeps=60)
    print("phase seconds: 60 substeps in batches of 20", flush=True)
    h = run_scene(gpu, sc, F, control_kw=dict(steps_per_launch=20))
    st = h["status"]
    ph = list(st.phase_seconds)
    assert st.launches == 3 and st.kernel_seconds > 0.0
    assert abs(sum(ph) - st.kernel_seconds) <= 1e-9 + 1e-6 * st.kernel_seconds
    for p in (3, 4, 5, 8, 10, 11, 12, 13):  # kick/classify, striker passes, sweep, velocity, sample, fold
        assert ph[p] > 0.0, p
