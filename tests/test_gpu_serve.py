"""The trial server: each answer is the same trial run alone, bit for bit.

Trials arrive staggered, wait for slots, run mixed with others that were
admitted earlier or later, and are answered as they stop. Every one is rerun
alone on the plain backend (its lattice, its marble, its run control) and
must give the same displacements, bond state, fragment labels and counters.
"""
import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.lattice import control
from banjo_amd.runner import prepare
from banjo_amd.serve import Trial, TrialServer, latency_summary

pytestmark = pytest.mark.gpu

TEMPLATES = [(8, 6, 8), (10, 5, 9)]
PLAN = [  # arrival s, template, speed, tilt, azimuth
    (0.000, 0, 2.5, 0.10, 0.3), (0.000, 1, 3.0, 0.00, 0.0), (0.010, 0, 0.8, 0.25, 2.0),
    (0.020, 1, 1.5, 0.20, 4.0), (0.050, 0, 3.0, 0.05, 1.0), (0.060, 1, 0.6, 0.30, 5.5)]


def trials():
    out = []
    for i, (t, k, v, tilt, az) in enumerate(PLAN):
        out.append(Trial(id=i, arrival=t, template=k, sphere=scenes.served_marble(TEMPLATES[k], v, tilt, az),
                         max_steps=360, calm_steps=120, quiet_steps=120, min_steps=120))
    return out


def test_served_trials_equal_alone(gpu):
    lats = [scenes.served_lattice(e) for e in TEMPLATES]
    server = TrialServer(gpu, lats, slots=3, chunk=16)
    try:
        answers = server.serve(trials(), keep_state=True)
    finally:
        server.close()
    assert sorted(a.id for a in answers) == list(range(len(PLAN)))
    by_id = {a.id: a for a in answers}
    for t in trials():
        a = by_id[t.id]
        sc = lats[t.template]
        sched, state = prepare(gpu, sc)
        be = gpu.backend(abi.PRECISION_FLOAT, sched, 0)
        be.upload(state, sc.settings, t.sphere)
        st = be.run(control(t.max_steps, calm_steps=t.calm_steps, quiet_steps=t.quiet_steps, min_steps=t.min_steps))
        out = state.copy()
        sp = be.download(out)
        labels, count = be.fragment_labels()
        be.close()
        assert np.array_equal(out.u, a.state.u), f"trial {t.id}"
        assert np.array_equal(out.v, a.state.v), f"trial {t.id}"
        assert np.array_equal(out.alive, a.state.alive), f"trial {t.id}"
        assert np.array_equal(out.damage, a.state.damage), f"trial {t.id}"
        assert np.array_equal(labels, a.labels) and count == a.fragments, f"trial {t.id}"
        assert st.total_steps == a.substeps and st.exit_reason == a.exit_reason, f"trial {t.id}"
        assert st.broken_bonds == a.broken_bonds, f"trial {t.id}"
        speed = float(np.sqrt(sum(x * x for x in sp.velocity)))
        assert (speed < 0.01) == a.settled


def test_admission_respects_the_memory_budget(gpu):
    lats = [scenes.served_lattice(e) for e in TEMPLATES]
    server = TrialServer(gpu, lats, slots=4, chunk=16)
    try:
        # Room for one trial of the larger template at a time.
        server.budget = max(server.footprint) + 1
        answers = server.serve(trials())
        assert server.peak_bytes <= server.budget
        assert server.peak_active == 1
        s = latency_summary(answers)
        assert s["trials"] == len(PLAN) and s["p99_s"] >= s["p50_s"] > 0
    finally:
        server.close()


def test_template_above_the_budget_is_refused(gpu):
    from banjo_amd.serve import TrialSession
    s = TrialSession(gpu, slots=2, max_templates=2, hbm_budget_bytes=1 << 20)
    try:
        with pytest.raises(ValueError):
            s.add_template(scenes.served_lattice((16, 8, 16)))
        s.budget = 1 << 30
        assert s.add_template(scenes.served_lattice((8, 6, 8))) == 0
    finally:
        s.close()
