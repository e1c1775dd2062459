"""Substep graphs against enqueued substeps, on the device.

A run replays a captured block of substeps (blx_tuning graph_substeps 8) or
enqueues every substep (the default). The two must give the same bits after every
run() call: calls that are whole graphs, graphs plus an enqueued remainder,
and calls shorter than one graph, with the host's synchronisations and
state reads between them. The same holds for a batch advanced in chunks.
"""
import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.lattice import control
from banjo_amd.runner import prepare, run_batch

pytestmark = pytest.mark.gpu
F = abi.PRECISION_FLOAT
CALLS = [8, 13, 3, 40, 16, 7]


def backend(gpu, sc, graphs):
    sched, state = prepare(gpu, sc)
    be = gpu.backend(F, sched, 0)
    be.upload(state, sc.settings, sc.sphere)
    be.set_tuning(graph_substeps=8 if graphs else 0)
    return be, state


@pytest.mark.parametrize("scene", ["impact", "bridge"])
def test_graph_replays_equal_enqueued_substeps(gpu, scene):
    sc = scenes.small_impact(nx=14, ny=8, nz=14, substeps=90) if scene == "impact" else scenes.bridge_impact()
    print(f"{scene}: calls of {CALLS} substeps, graphs against enqueued", flush=True)
    g, state = backend(gpu, sc, True)
    e, _ = backend(gpu, sc, False)
    try:
        for n in CALLS:
            sg, se = g.run(control(n)), e.run(control(n))
            assert sg.total_steps == se.total_steps and sg.broken_bonds == se.broken_bonds
            assert sg.removed_energy_j == se.removed_energy_j and sg.max_damage == se.max_damage
            assert sg.contact.impulse_contacts == se.contact.impulse_contacts
            a, b = state.copy(), state.copy()
            sa, sb = g.download(a), e.download(b)
            for name in ("u", "v", "u_prev", "alive", "damage", "failure_mode", "prev_tensile", "prev_compressive",
                         "prev_shear"):
                assert np.array_equal(getattr(a, name), getattr(b, name)), (n, name)
            assert list(sa.center) == list(sb.center) and list(sa.velocity) == list(sb.velocity)
        la, ca = g.fragment_labels()
        lb, cb = e.fragment_labels()
        assert ca == cb and np.array_equal(la, lb)
        assert sg.broken_bonds > 0 and sg.contact.impulse_contacts > 0
    finally:
        g.close()
        e.close()


def test_batch_graphs_equal_enqueued_substeps(gpu):
    trials = [scenes.config3_trial(t, substeps=96) for t in range(4)]
    eager = run_batch(gpu, trials, F, steps_per_launch=24)
    graphs = run_batch(gpu, trials, F, steps_per_launch=24, tuning={"graph_substeps": 8})
    for a, b in zip(graphs, eager):
        for k in ("u", "v", "alive", "damage", "labels"):
            assert np.array_equal(a[k], b[k]), k
        assert a["status"].broken_bonds == b["status"].broken_bonds


def test_batch_graphs_with_uneven_budgets_equal_enqueued(gpu):
    # Trials stop at different substeps, so the launch maps are rebuilt between
    # graph replays; a rebuild must drop the captured graph (its launch
    # arguments hold the old maps) and capture again.
    trials = [scenes.config3_trial(t, substeps=96) for t in range(4)]
    steps = [40, 96, 17, 64]
    kws = [{}, {"no_failure_steps": 30}, {}, {"calm_steps": 20}]
    print(f"batch of {len(trials)} with budgets {steps}, graphs against enqueued", flush=True)
    eager = run_batch(gpu, trials, F, substeps=steps, control_kw=kws, steps_per_launch=8)
    graphs = run_batch(gpu, trials, F, substeps=steps, control_kw=kws, steps_per_launch=8,
                       tuning={"graph_substeps": 8})
    for a, b in zip(graphs, eager):
        for k in ("u", "v", "alive", "damage", "labels"):
            assert np.array_equal(a[k], b[k]), k
        assert a["status"].total_steps == b["status"].total_steps
        assert a["status"].exit_reason == b["status"].exit_reason
        assert a["status"].broken_bonds == b["status"].broken_bonds
