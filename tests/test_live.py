"""Fracture jobs run by worker threads (banjo_amd/live.py), the live world's
work step (src/fastlattice/LiveWorld.cpp:12727-12753), and their JSON-lines
wire (banjo_amd/live_wire.py), here over the oracle library on the host:

- jobs run together on worker threads, each on its own backend, give what
  the same lattice gives run alone, bit for bit, with the labels and the
  fragment mass properties of that run;
- RunControl::abandon reaches the job's run (the serial lane the oracle
  restates reads no flag; the device's run stops at its next batch);
- the wire keeps the engine process's transport (tools/live_world_run.cpp:
  1-10, 2115-2125, 3558-3566), reports breaks that finished on a later
  "step", and its poses conserve the body's mass and momentum.

tests/test_gpu_live.py runs the same jobs on the device.
"""
import io
import json
import time

import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.live import EXIT_ABANDONED, FractureJob, LiveWork
from banjo_amd.live_wire import LiveWire
from banjo_amd.serve_wire import serve_lines
from banjo_amd.runner import run_scene
from test_fragment_mass import lists

F = abi.PRECISION_FLOAT


def jobs():
    a = scenes.served_lattice((10, 4, 9), seed=5, name="slab")
    a.sphere = scenes.served_marble((10, 4, 9), 2.5, 0.1, 0.7)
    b = scenes.small_impact(nx=10, ny=6, nz=10, substeps=60)
    c = scenes.served_lattice((8, 5, 8), seed=9, name="block")
    c.sphere = scenes.served_marble((8, 5, 8), 3.0, 0.2, 2.0)
    return [FractureJob("slab", a, substeps=90), FractureJob("impact", b), FractureJob("block", c, substeps=70)]


def alone(lib, job):
    sc = job.scene
    res = run_scene(lib, sc, F, substeps=job.substeps)
    off, nodes = lists(res["labels"])
    st = res["state"]
    pos = (st.origin[None, :] + st.x0.reshape(-1, 3)) + st.u.reshape(-1, 3)
    return res, lib.fragment_mass(pos, st.v.reshape(-1, 3), st.mass, sc.cell, off, nodes)


def test_jobs_on_workers_equal_runs_alone(oracle):
    work = LiveWork(oracle, workers=2)
    try:
        js = jobs()
        ids = [work.submit(j) for j in js]
        got = {r.job: r for r in (work.wait(k) for k in ids)}
    finally:
        work.close()
    for k, job in zip(ids, js):
        r = got[k]
        want, mass = alone(oracle, job)
        print(job.name, r.status.total_steps, r.status.broken_bonds, r.fragments, f"{r.cost_ms:.0f} ms")
        assert r.name == job.name and not r.abandoned
        assert np.array_equal(r.state.u, want["state"].u)
        assert np.array_equal(r.state.v, want["state"].v)
        assert np.array_equal(r.state.alive, want["state"].alive)
        assert r.status.total_steps == want["status"].total_steps
        assert r.status.broken_bonds == want["status"].broken_bonds
        assert r.status.exit_reason == want["status"].exit_reason
        assert np.array_equal(r.labels, want["labels"]) and r.fragments == want["components"]
        assert r.mass.tobytes() == mass.tobytes()
    assert any(got[k].status.broken_bonds > 0 for k in ids)


def test_poll_returns_each_job_once(oracle):
    work = LiveWork(oracle, workers=2)
    try:
        ids = [work.submit(j) for j in jobs()]
        seen = []
        deadline = time.time() + 120
        while len(seen) < len(ids) and time.time() < deadline:
            seen += [r.job for r in work.poll()]
            time.sleep(0.01)
        assert sorted(seen) == sorted(ids)
        assert work.poll() == [] and work.running() == 0
    finally:
        work.close()


def test_abandon_reaches_the_run(oracle):
    # One worker: the second job waits behind the first, so it is abandoned
    # before its run starts. The oracle restates the reference's serial lane,
    # which reads no abandon flag and runs to its end (FastLattice.hpp:122-127:
    # only the parallel CPU lane reads it); tests/test_gpu_live.py checks that
    # the device's run stops at its next batch.
    work = LiveWork(oracle, workers=1)
    try:
        first = work.submit(jobs()[1])
        sc = scenes.small_impact(nx=10, ny=6, nz=10, substeps=200)
        second = work.submit(FractureJob("long", sc, steps_per_launch=8))
        work.abandon(second)
        r = work.wait(second)
        assert work.wait(first).status.exit_reason != EXIT_ABANDONED
    finally:
        work.close()
    want = run_scene(oracle, sc, F)
    assert not r.abandoned and r.status.total_steps == want["status"].total_steps
    assert np.array_equal(r.state.u, want["state"].u)
    with pytest.raises(KeyError):
        work.abandon(12345)


def talk(lib, lines, workers=2):
    work = LiveWork(lib, workers=workers)
    out = io.StringIO()
    try:
        serve_lines(LiveWire(work), [json.dumps(x) if not isinstance(x, str) else x for x in lines], out)
    finally:
        work.close()
    return [json.loads(line) for line in out.getvalue().splitlines()]


def test_wire_transport_and_breaks(oracle):
    replies = talk(oracle, [
        {"op": "body", "name": "pane", "edges": [10, 3, 10], "seed": 7},
        "",
        {"op": "body", "name": "pane", "edges": [4, 4, 4]},
        {"op": "fracture", "name": "nothing"},
        "not json",
        {"op": "fracture", "name": "pane", "speed_m_s": 3.0, "tilt_rad": 0.1, "azimuth_rad": 1.0, "max_steps": 120},
        {"op": "poses", "max_pieces": 100000},
        {"op": "frobnicate"},
        {"op": "quit"},
        {"op": "body", "name": "after", "edges": [4, 4, 4]},
    ])
    opening, body, dup, missing, bad, done, poses, unknown, bye = replies
    assert opening["ok"] and opening["schema"] == "banjo.live.v1" and opening["workers"] == 2
    assert body == {"ok": True, "body": "pane", "nodes": 300, "bonds": body["bonds"]} and body["bonds"] > 0
    for r in (dup, missing, bad, unknown):
        assert r["ok"] is False and r["error"]
    assert done["ok"] and done["fracture"] == "pane" and done["result"]["substeps"] <= 120
    assert done["result"]["fragments"] == len(poses["pieces"])
    assert bye == {"ok": True, "bye": True}
    # The pieces are the body, split: mass and momentum add up to the body's.
    sc = scenes.served_lattice((10, 3, 10), seed=7)
    sc.sphere = scenes.served_marble((10, 3, 10), 3.0, 0.1, 1.0)
    want = run_scene(oracle, sc, F, substeps=120, control_kw=dict(calm_steps=480, quiet_steps=480, min_steps=480))
    st = want["state"]
    m = np.array([p["mass_kg"] for p in poses["pieces"]])
    v = np.array([p["velocity_m_s"] for p in poses["pieces"]])
    assert m.sum() == pytest.approx(st.mass.sum(), rel=1e-12)
    np.testing.assert_allclose((m[:, None] * v).sum(0), (st.mass[:, None] * st.v.reshape(-1, 3)).sum(0),
                               rtol=1e-9, atol=1e-12)
    assert done["result"]["broken_bonds"] == want["status"].broken_bonds


def test_wire_reports_pending_breaks_on_step(oracle):
    lines = [{"op": "body", "name": "a", "edges": [8, 4, 8]}, {"op": "body", "name": "b", "edges": [9, 3, 7]},
             {"op": "fracture", "name": "a", "wait": False, "speed_m_s": 2.0, "max_steps": 60},
             {"op": "fracture", "name": "b", "wait": False, "speed_m_s": 2.5, "max_steps": 60}]
    # Enough steps for both breaks to finish on the host (each takes well under a second).
    lines += [{"op": "step", "n": 1}] * 400 + [{"op": "poses"}, {"op": "quit"}]
    work = LiveWork(oracle, workers=2)
    wire = LiveWire(work)
    out = []
    try:
        out.append(wire.opening())
        for cmd in lines:
            out.append(wire.handle(cmd))
            if cmd["op"] == "step" and sum(len(r.get("fractures", [])) for r in out) == 2:
                break
            if cmd["op"] == "step":
                time.sleep(0.01)
        out.append(wire.handle({"op": "poses"}))
    finally:
        work.close()
    assert out[3] == {"ok": True, "fracture": "a", "job": 1, "pending": True}
    assert out[4] == {"ok": True, "fracture": "b", "job": 2, "pending": True}
    finished = [f for r in out if "fractures" in r for f in r["fractures"]]
    assert sorted((f["job"], f["body"]) for f in finished) == [(1, "a"), (2, "b")]
    assert {p["body"] for p in out[-1]["pieces"]} == {"a", "b"}
