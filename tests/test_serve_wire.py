"""The JSON-lines wire of the served form (banjo_amd/serve_wire.py): its
transport, as the reference's engine process speaks it
(tools/live_world_run.cpp:1-10, 2115-2125, 3558-3566): an opening reply,
one reply line per request, {"ok":false,"error":...} for a failed request
with the session going on, empty lines skipped, "quit" ending the session.
The device side is replaced here by a stand-in session that answers every
trial after two chunks; tests/test_gpu_serve_wire.py runs the real one."""
import io
import json

from banjo_amd import scenes
from banjo_amd.serve import Answer
from banjo_amd.serve_wire import Wire, serve_lines


class StandIn:
    """The TrialSession interface, answering each trial two pumps after it was submitted."""
    slots, max_templates, budget, chunk = 2, 4, 1 << 30, 16

    def __init__(self):
        self.templates, self.footprint, self.pending, self.active = [], [], [], {}
        self.t, self.peak_active, self.peak_bytes = 0.0, 0, 0

    def now(self):
        return self.t

    def add_template(self, sc):
        self.templates.append(sc)
        self.footprint.append(1000 * sc.node_count)
        return len(self.templates) - 1

    def submit(self, t):
        self.pending.append(t)

    @property
    def busy(self):
        return bool(self.pending or self.active)

    def pump(self, keep_state=False):
        self.t += 0.5
        while self.pending and len(self.active) < self.slots:
            t = self.pending.pop(0)
            self.active[t.id] = [t, self.t, 2]
        done = []
        for k in list(self.active):
            self.active[k][2] -= 1
            if self.active[k][2] == 0:
                t, admitted, _ = self.active.pop(k)
                done.append(Answer(id=t.id, arrival=t.arrival, admitted=admitted, finished=self.t, fragments=1,
                                   settled=True, exit_reason=1, substeps=t.max_steps, broken_bonds=0, slot=0))
        self.peak_active = max(self.peak_active, len(self.active))
        return done


def talk(lines):
    out = io.StringIO()
    serve_lines(Wire(StandIn()), [json.dumps(x) if not isinstance(x, str) else x for x in lines], out)
    return [json.loads(line) for line in out.getvalue().splitlines()]


def test_wire_transport_and_session():
    replies = talk([
        {"op": "template", "name": "small", "edges": [6, 5, 6]},
        "",
        {"op": "trial", "id": "a", "template": "small", "speed_m_s": 2.0, "tilt_rad": 0.1},
        {"op": "trial", "id": "b", "template": "small", "max_steps": 240,
         "sphere": {"center_m": [0, 0.2, 0], "velocity_m_s": [0, -1, 0], "radius_m": 0.05, "mass_kg": 0.05}},
        {"op": "trial", "id": "c", "template": "small"},
        "{not json",
        {"op": "trial", "template": "missing"},
        {"op": "fly"},
        {"op": "step", "n": 2},
        {"op": "drain"},
        {"op": "stats"},
        {"op": "quit"},
        {"op": "stats"},  # after quit: never read
    ])
    opening, tpl, qa, qb, qc, bad_json, missing, unknown, step, drain, stats, bye = replies
    assert opening["ok"] and opening["engine"] == "banjo_amd" and opening["dt_s"] == scenes.SERVE_DT
    sc = scenes.served_lattice((6, 5, 6))
    assert tpl == {"ok": True, "template": "small", "index": 0, "nodes": sc.node_count, "bonds": sc.bond_count,
                   "bytes": 1000 * sc.node_count}
    assert qa["ok"] and qa["queued"] == "a" and qc["pending"] == 3
    for r in (bad_json, missing, unknown):
        assert r["ok"] is False and r["error"]
    assert "missing" in missing["error"] and "fly" in unknown["error"]
    assert [a["id"] for a in step["answers"]] == ["a", "b"] and step["pending"] == 1 and step["running"] == 0
    assert step["answers"][1]["substeps"] == 240 and step["answers"][0]["substeps"] == scenes.SERVE_SUBSTEPS
    assert [a["id"] for a in drain["answers"]] == ["c"] and drain["running"] == 0
    assert stats["answered"] == 3 and stats["p99_s"] >= stats["p50_s"] > 0
    assert bye == {"ok": True, "bye": True}


def test_wire_refuses_bad_requests():
    replies = talk([
        {"op": "template", "name": "t", "edges": [6, 5]},
        {"op": "template", "name": "t", "edges": [6, 5, 6]},
        {"op": "template", "name": "t", "edges": [6, 5, 6]},
        {"op": "trial", "template": "t", "max_steps": 0},
        {"op": "trial", "template": "t", "sphere": {"center_m": [0, "x", 0]}},
        {"op": "trial", "template": "t", "sphere": {"center_m": [0, 1, 0], "radius_m": -1}},
        {"op": "step", "n": 0},
        [1, 2],
    ])
    oks = [r["ok"] for r in replies[1:]]
    assert oks == [False, True, False, False, False, False, False, False]
