"""The batched, served and cut workloads against the oracle, at their sizes.

test_gpu_batch.py and test_gpu_serve.py hold every batched or served lattice
to the same lattice run alone on the plain HIP backend. Here the forms are held
to the oracle directly, on lattices of the workloads' own size distributions:

- batched: eight cantilevers drawn from the batch's edges (log-uniform in
  [8, 48] per axis), the smallest and the largest included, advanced together
  in one batch with per-lattice budgets; each equals the oracle's run of it,
  the smallest over all 1,920 substeps of the workload, the largest over its
  first 60;
- served: a seeded stream of marble trials over lattices of the same
  distribution, admitted as they arrive into a running batch while others are
  mid-run; each answer (displacements, velocities, bonds, labels, exit reason,
  substeps, settled flag) equals the oracle's run of that trial alone;
- cut: the 256x32x32 plastic plank under the 1 mm blade, its first 36
  substeps, in which the blade has engaged and bonds have broken.
"""
import math

import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.lattice import control
from banjo_amd.runner import run_batch, run_scene
from banjo_amd.serve import SETTLED_SPEED, Trial, TrialServer
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu
F = abi.PRECISION_FLOAT

# (trial index in batch_edges(971, 512), substeps): the three smallest lattices
# of the distribution run the workload's whole 1,920 substeps; the quartiles
# and the largest (45x47x44, 93,060 nodes) run shorter windows.
BATCH_PICK = [(310, 1920), (439, 1920), (215, 1920), (172, 960), (86, 480), (335, 240), (465, 120), (311, 60)]


def test_config3_batch_equals_the_oracle(gpu, oracle):
    edges = scenes.batch_edges(971, 512)
    scs = [scenes.config3_trial(t, edges=edges[t], substeps=n) for t, n in BATCH_PICK]
    print("config 3 batch: " + ", ".join(f"{sc.name} x{n}" for sc, (_, n) in zip(scs, BATCH_PICK)), flush=True)
    batched = run_batch(gpu, scs, F, substeps=[n for _, n in BATCH_PICK])
    broken = 0
    for sc, b in zip(scs, batched):
        print(f"  oracle {sc.name}", flush=True)
        o = run_scene(oracle, sc, F)
        assert b["status"].total_steps == sc.substeps
        # The load ledger is the one sum the device takes in a fixed tree order
        # rather than node order. Summing N same-sign terms in two orders can
        # differ by up to about N unit roundoffs relative (the recursive
        # summation bound), so its tolerance grows with the lattice: 2e-11 for
        # the largest, where 1.75e-12 was measured. Everything else is exact.
        assert_same(b, o, ledger_rtol=max(1e-12, sc.node_count * 2.0 ** -52))
        broken += b["status"].broken_bonds
    assert broken > 0


# Served templates from the stream's own pool (batch_edges(4242, 64)): the
# smallest, a median and a large lattice.
SERVE_POOL = [31, 26, 48]
SERVE_PLAN = [  # arrival s, template, speed, tilt, azimuth, max substeps
    (0.000, 0, 2.5, 0.10, 0.3, 480), (0.000, 1, 3.0, 0.00, 0.0, 240), (0.005, 2, 1.2, 0.20, 1.0, 120),
    (0.010, 0, 0.8, 0.25, 2.0, 480), (0.015, 1, 1.5, 0.20, 4.0, 240), (0.020, 0, 3.0, 0.05, 1.0, 480),
    (0.030, 2, 2.8, 0.15, 5.0, 120), (0.040, 1, 0.6, 0.30, 5.5, 240)]
SERVE_RULES = dict(calm_steps=60, quiet_steps=60, min_steps=60)


def test_config4_served_trials_equal_the_oracle(gpu, oracle):
    pool = scenes.batch_edges(4242, 64)
    edges = [pool[k] for k in SERVE_POOL]
    lats = [scenes.served_lattice(e, seed=4242 + k) for k, e in zip(SERVE_POOL, edges)]
    trials = [Trial(id=i, arrival=a, template=k, sphere=scenes.served_marble(edges[k], v, tilt, az), max_steps=n,
                    **SERVE_RULES)
              for i, (a, k, v, tilt, az, n) in enumerate(SERVE_PLAN)]
    print("config 4 stream: templates " + ", ".join(sc.name for sc in lats), flush=True)
    server = TrialServer(gpu, lats, slots=3, chunk=16)
    try:
        answers = server.serve(trials, keep_state=True)
    finally:
        server.close()
    assert sorted(a.id for a in answers) == list(range(len(SERVE_PLAN)))
    assert server.peak_active == 3  # trials overlapped on the device
    by_id = {a.id: a for a in answers}
    reasons = set()
    for t in trials:
        a = by_id[t.id]
        sc = lats[t.template]
        sc.sphere = t.sphere
        print(f"  oracle trial {t.id} on {sc.name}", flush=True)
        o = run_scene(oracle, sc, F, substeps=t.max_steps, control_kw=SERVE_RULES)
        st = o["status"]
        assert np.array_equal(o["state"].u, a.state.u), f"trial {t.id}"
        assert np.array_equal(o["state"].v, a.state.v), f"trial {t.id}"
        assert np.array_equal(o["state"].alive, a.state.alive), f"trial {t.id}"
        assert np.array_equal(o["state"].damage, a.state.damage), f"trial {t.id}"
        assert np.array_equal(o["labels"], a.labels) and o["components"] == a.fragments, f"trial {t.id}"
        assert st.total_steps == a.substeps and st.exit_reason == a.exit_reason, f"trial {t.id}"
        assert st.broken_bonds == a.broken_bonds, f"trial {t.id}"
        speed = math.sqrt(sum(x * x for x in o["sphere"].velocity))
        assert (speed < SETTLED_SPEED) == a.settled, f"trial {t.id}"
        reasons.add(a.exit_reason)
    assert 3 in reasons and len(reasons) > 1  # budgets and early exits both answered


def test_config5_cut_first_substeps_equal_the_oracle(gpu, oracle):
    sc = scenes.config5(substeps=36)
    print(f"config 5 N {sc.node_count} B {sc.bond_count}, 36 substeps, HIP against the oracle", flush=True)
    h = run_scene(gpu, sc, F, labels="device")
    o = run_scene(oracle, sc, F)
    assert h["status"].broken_bonds > 0 and h["status"].plastic_work_j > 0.0
    assert_same(h, o)
