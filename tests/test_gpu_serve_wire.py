"""The JSON-lines wire on the device: trials sent as request lines are
answered with what the same trial gives alone on the plain backend (its
lattice, its marble, its step budget), and the served session's refactor
keeps the trial server's answers."""
import io
import json

import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.lattice import control
from banjo_amd.runner import prepare
from banjo_amd.serve import TrialSession
from banjo_amd.serve_wire import Wire, serve_lines

pytestmark = pytest.mark.gpu

TEMPLATES = {"a": (8, 6, 8), "b": (10, 5, 9)}
TRIALS = [("t0", "a", 2.5, 0.10, 0.3), ("t1", "b", 3.0, 0.0, 0.0), ("t2", "a", 0.8, 0.25, 2.0),
          ("t3", "b", 1.5, 0.20, 4.0), ("t4", "a", 3.0, 0.05, 1.0)]
STEPS = 360


def test_wire_answers_equal_alone(gpu):
    lines = [json.dumps({"op": "template", "name": n, "edges": list(e)}) for n, e in TEMPLATES.items()]
    lines += [json.dumps({"op": "trial", "id": i, "template": t, "speed_m_s": v, "tilt_rad": tilt,
                          "azimuth_rad": az, "max_steps": STEPS}) for i, t, v, tilt, az in TRIALS[:3]]
    lines += [json.dumps({"op": "step", "n": 3})]
    lines += [json.dumps({"op": "trial", "id": i, "template": t, "speed_m_s": v, "tilt_rad": tilt,
                          "azimuth_rad": az, "max_steps": STEPS}) for i, t, v, tilt, az in TRIALS[3:]]
    lines += [json.dumps({"op": "drain"}), json.dumps({"op": "stats"}), json.dumps({"op": "quit"})]
    session = TrialSession(gpu, slots=2, max_templates=2, chunk=32)
    out = io.StringIO()
    try:
        serve_lines(Wire(session), lines, out)
    finally:
        session.close()
    replies = [json.loads(x) for x in out.getvalue().splitlines()]
    assert all(r["ok"] for r in replies), [r for r in replies if not r["ok"]]
    answers = {a["id"]: a for r in replies for a in r.get("answers", [])}
    assert sorted(answers) == [t[0] for t in TRIALS]
    assert replies[-2]["answered"] == len(TRIALS) and replies[-2]["peak_running"] <= 2
    for i, t, v, tilt, az in TRIALS:
        sc = scenes.served_lattice(TEMPLATES[t])
        sched, state = prepare(gpu, sc)
        be = gpu.backend(abi.PRECISION_FLOAT, sched, 0)
        be.upload(state, sc.settings, scenes.served_marble(TEMPLATES[t], v, tilt, az))
        st = be.run(control(STEPS, calm_steps=480, quiet_steps=480, min_steps=480))
        sp = be.download(state.copy())
        _, count = be.fragment_labels()
        be.close()
        a = answers[i]
        assert (a["substeps"], a["exit_reason"], a["broken_bonds"], a["fragments"]) == \
            (st.total_steps, st.exit_reason, st.broken_bonds, count), i
        assert a["settled"] == (float(np.sqrt(sum(x * x for x in sp.velocity))) < 0.01), i
