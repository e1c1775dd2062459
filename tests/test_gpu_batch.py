"""Batches of lattices on the HIP engine against the same lattices run alone.

A batch advances independent lattices of different sizes in shared launches
(blx_batch_* of include/banjo_lattice.h). Each lattice's results must be
those of the same lattice on its own backend, bit for bit: displacements,
velocities, bond state, labels and the exact status counters. The plain
backend is itself checked against the oracle in test_gpu_parity.py, so the
batched form inherits that pin.
"""
import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.lattice import Batch
from banjo_amd.runner import prepare, run_batch, run_scene
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu

EDGES = [(6, 5, 4), (9, 7, 5), (12, 6, 8), (5, 11, 6), (7, 7, 7)]


def cantilevers(substeps=120):
    return [scenes.config3_trial(t, edges=e, substeps=substeps) for t, e in enumerate(EDGES)]


def alone(gpu, sc, substeps=None, control_kw=None):
    return run_scene(gpu, sc, substeps=substeps, labels="device", control_kw=control_kw)


def test_batch_equals_alone(gpu):
    scs = cantilevers()
    batched = run_batch(gpu, scs)
    for sc, b in zip(scs, batched):
        a = alone(gpu, sc)
        assert a["status"].broken_bonds == b["status"].broken_bonds
        assert_same(a, b)


def test_batch_breaks_bonds(gpu):
    # The tip loads break bonds in most of these bars, so the batched folds run.
    batched = run_batch(gpu, cantilevers(240))
    assert sum(b["status"].broken_bonds > 0 for b in batched) >= 2


def test_batch_uneven_budgets_and_exit_rules(gpu):
    scs = cantilevers()
    steps = [60, 120, 90, 30, 150]
    kws = [{}, {}, {"no_failure_steps": 40}, {}, {"calm_steps": 20, "calm_damage_margin": 0.5}]
    batched = run_batch(gpu, scs, substeps=steps, control_kw=kws)
    for sc, n, kw, b in zip(scs, steps, kws, batched):
        a = alone(gpu, sc, substeps=n, control_kw=kw)
        assert a["status"].exit_reason == b["status"].exit_reason
        assert_same(a, b)


def test_exit_reasons_in_one_call_equal_each_status(gpu):
    # blx_batch_exit_reasons: 0 for a slot still running or never filled, else
    # the exit reason blx_batch_status reports.
    from banjo_amd.runner import make_control
    scs = cantilevers()[:3]
    bt = Batch(gpu, abi.PRECISION_FLOAT, 4)
    try:
        for k, (sc, n) in enumerate(zip(scs, [30, 60, 90])):
            sched, state = prepare(gpu, sc)
            bt.upload(k, sched, state, sc.settings, sc.sphere)
            bt.arm(k, make_control(n))
        bt.advance(45)
        got = bt.exit_reasons(4)
        assert list(got) == [bt.status(k).exit_reason for k in range(3)] + [0]
        assert got[0] != 0  # its 30-substep budget has run out
        while bt.advance(1 << 62):
            pass
        got = bt.exit_reasons(3)
        assert list(got) == [bt.status(k).exit_reason for k in range(3)] and all(got != 0)
        with pytest.raises(ValueError):
            bt.exit_reasons(5)
    finally:
        bt.close()


def test_batch_impacts_overlap_the_striker_pass(gpu):
    # Same height (so the same ground plane), different footprints, speeds and seeds.
    scs = [scenes.impact(nx, 8, nz, ball_radius=0.03, ball_mass=0.5, ball_speed=v, substeps=60, seed=s,
                         name=f"impact_{nx}x8x{nz}")
           for nx, nz, v, s in [(12, 12, 6.0, 11), (16, 10, 8.0, 12), (10, 14, 4.0, 13)]]
    batched = run_batch(gpu, scs)
    for sc, b in zip(scs, batched):
        a = alone(gpu, sc)
        assert a["status"].contact.impulse_contacts == b["status"].contact.impulse_contacts
        assert_same(a, b)


def test_released_slot_takes_a_new_lattice(gpu):
    scs = cantilevers(80)
    first, second, smaller = scs[1], scs[2], scs[0]
    bt = Batch(gpu, abi.PRECISION_FLOAT, 2)
    try:
        for k, sc in enumerate((first, second)):
            sched, state = prepare(gpu, sc)
            bt.upload(k, sched, state, sc.settings, sc.sphere)
            bt.set_external_forces(k, sc.forces, sc.force_substeps)
            bt.arm(k, _control(80))
        while bt.advance(1 << 40):
            pass
        bytes_before = bt.slot_bytes(0)
        bt.release(0)
        # A smaller lattice in the released slot reuses its memory.
        sched, state = prepare(gpu, smaller)
        bt.upload(0, sched, state, smaller.settings, smaller.sphere)
        bt.set_external_forces(0, smaller.forces, smaller.force_substeps)
        bt.arm(0, _control(80))
        while bt.advance(1 << 40):
            pass
        assert bt.slot_bytes(0) == bytes_before
        out = state.copy()
        bt.download(0, out)
        a = alone(gpu, smaller, substeps=80)
        assert np.array_equal(out.u.reshape(-1, 3), a["u"])
        assert np.array_equal(out.v.reshape(-1, 3), a["v"])
        assert bt.status(0).broken_bonds == a["status"].broken_bonds
    finally:
        bt.close()


def test_batch_refuses_mixed_settings(gpu):
    a = scenes.config3_trial(0, edges=(5, 5, 5), substeps=10)
    b = scenes.config3_trial(1, edges=(6, 5, 5), substeps=10)
    b.settings.dt = a.settings.dt * 0.5
    bt = Batch(gpu, abi.PRECISION_FLOAT, 2)
    try:
        s, st = prepare(gpu, a)
        bt.upload(0, s, st, a.settings, a.sphere)
        s, st = prepare(gpu, b)
        with pytest.raises(ValueError):
            bt.upload(1, s, st, b.settings, b.sphere)
    finally:
        bt.close()


def test_batch_in_double(gpu):
    scs = cantilevers(60)[:3]
    batched = run_batch(gpu, scs, precision=abi.PRECISION_DOUBLE)
    for sc, b in zip(scs, batched):
        a = run_scene(gpu, sc, precision=abi.PRECISION_DOUBLE, labels="device")
        assert_same(a, b)


def _control(n):
    from banjo_amd.lattice import control
    return control(n)


def test_batch_blade_cuts_equal_alone(gpu):
    # Same length and height (so the same ledges and blade), different widths and seeds.
    scs = [scenes.plank_cut(20, 6, nz, speed=4.0, substeps=90, ledge_cells=4, seed=s, name=f"plank_20x6x{nz}")
           for nz, s in [(4, 31), (6, 32), (9, 33)]]
    batched = run_batch(gpu, scs)
    for sc, b in zip(scs, batched):
        a = alone(gpu, sc)
        assert a["status"].broken_bonds > 0
        assert_same(a, b)
