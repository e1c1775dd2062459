"""Fracture jobs on the device (banjo_amd/live.py): several jobs run by
worker threads at once, each on its own backend and HIP streams, give what
the same lattice gives run alone on the device, bit for bit, including the
device's fragment labels and mass properties; RunControl::abandon stops a
device run at its next batch (exit reason 7, FastLattice.hpp:122-127); the
wire (banjo_amd/live_wire.py) answers breaks that finish while it goes on.
"""
import time

import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.live import EXIT_ABANDONED, FractureJob, LiveWork
from banjo_amd.live_wire import LiveWire
from banjo_amd.runner import run_scene
from test_live import jobs

pytestmark = pytest.mark.gpu
F = abi.PRECISION_FLOAT


def test_device_jobs_equal_runs_alone(gpu):
    js = jobs() + [FractureJob("wide", scenes.small_impact(nx=24, ny=8, nz=24, substeps=80))]
    work = LiveWork(gpu, workers=3)
    try:
        ids = [work.submit(j) for j in js]
        got = {k: work.wait(k, timeout=300) for k in ids}
    finally:
        work.close()
    for k, job in zip(ids, js):
        r = got[k]
        want = run_scene(gpu, job.scene, F, substeps=job.substeps, labels="device", keep_backend=True)
        be = want["backend"]
        try:
            mass = be.fragment_mass_properties(job.scene.cell)
        finally:
            be.close()
        print(job.name, r.status.total_steps, r.status.broken_bonds, r.fragments, f"{r.cost_ms:.0f} ms", flush=True)
        assert not r.abandoned
        for name in ("u", "v", "alive", "damage"):
            assert np.array_equal(getattr(r.state, name), getattr(want["state"], name)), name
        assert r.status.total_steps == want["status"].total_steps
        assert r.status.broken_bonds == want["status"].broken_bonds
        assert np.array_equal(r.labels, want["labels"]) and r.fragments == want["components"]
        assert r.mass.tobytes() == mass.tobytes()
    assert any(got[k].status.broken_bonds > 0 for k in ids)


def test_abandon_stops_a_device_run(gpu):
    work = LiveWork(gpu, workers=1)
    try:
        first = work.submit(jobs()[1])
        sc = scenes.small_impact(nx=10, ny=6, nz=10, substeps=400)
        second = work.submit(FractureJob("long", sc, steps_per_launch=8))
        work.abandon(second)
        r = work.wait(second, timeout=300)
        assert work.wait(first, timeout=300).status.exit_reason != EXIT_ABANDONED
    finally:
        work.close()
    assert r.abandoned and r.status.exit_reason == EXIT_ABANDONED
    assert r.status.total_steps == 8


def test_wire_on_the_device(gpu):
    work = LiveWork(gpu, workers=2)
    wire = LiveWire(work)
    try:
        assert wire.handle({"op": "body", "name": "a", "edges": [12, 4, 12]})["ok"]
        assert wire.handle({"op": "body", "name": "b", "edges": [10, 5, 9]})["ok"]
        assert wire.handle({"op": "fracture", "name": "a", "wait": False, "speed_m_s": 3.0})["pending"]
        assert wire.handle({"op": "fracture", "name": "b", "wait": False, "speed_m_s": 2.0})["pending"]
        finished = []
        deadline = time.time() + 300
        while len(finished) < 2 and time.time() < deadline:
            finished += wire.handle({"op": "step", "n": 1})["fractures"]
            time.sleep(0.005)
        poses = wire.handle({"op": "poses", "max_pieces": 1 << 20})
    finally:
        work.close()
    assert sorted(f["body"] for f in finished) == ["a", "b"]
    for f in finished:
        edges = {"a": (12, 4, 12), "b": (10, 5, 9)}[f["body"]]
        pieces = [p for p in poses["pieces"] if p["body"] == f["body"]]
        assert len(pieces) == f["fragments"]
        assert sum(p["nodes"] for p in pieces) == edges[0] * edges[1] * edges[2]
