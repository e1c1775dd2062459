"""Fracture jobs over JSON lines: the engine's wire for the lattice lane.

The reference's engine process holds a live world open and speaks JSON over
stdin and stdout, one object per line: an opening reply, then one reply per
request, ``{"ok":false,"error":...}`` for a request that failed, empty lines
skipped (tools/live_world_run.cpp:1-10, 2115-2125, 3558-3566). Of its ops,
those of the lattice lane are ``fracture`` (break a body; ``"wait":false``
goes on while the break runs, :17), ``step`` (the reply carries what has
finished), ``poses`` and ``quit``. This process speaks them over lattice
bodies, each break a banjo_amd.live job run on the device by a worker thread
while the wire goes on answering:

    python -m banjo_amd.live_wire [--workers 2]

  in   {"op":"body","name":"pane","edges":[24,4,24],"seed":971}
       {"op":"fracture","name":"pane","wait":false,"speed_m_s":3.0,"tilt_rad":0.1,"azimuth_rad":0.5}
       {"op":"fracture","name":"pane","sphere":{"center_m":[0,0.1,0],"velocity_m_s":[0,-2,0],
        "radius_m":0.05,"mass_kg":0.05},"max_steps":960}
       {"op":"step","n":1}        the breaks finished since the last reply
       {"op":"abandon","job":2}   RunControl::abandon for a break still running
       {"op":"poses","max_pieces":64}   the pieces of each body's last finished break
       {"op":"quit"}
  out  {"ok":true,"engine":"banjo_amd","schema":"banjo.live.v1","workers":2,...}
       {"ok":true,"body":"pane","nodes":2304,"bonds":29742}
       {"ok":true,"fracture":"pane","job":1,"pending":true}
       {"ok":true,"fracture":"pane","job":1,"result":{...}}      (waited for)
       {"ok":true,"running":1,"fractures":[{"job":1,"body":"pane","exit_reason":1,"substeps":960,
        "broken_bonds":812,"fragments":5,"abandoned":false,"cost_ms":41.2}]}
       {"ok":true,"pieces":[{"body":"pane","job":1,"piece":0,"nodes":2290,"mass_kg":2.29,
        "center_of_mass_m":[...],"velocity_m_s":[...],"angular_velocity_rad_s":[...]}, ...]}

Bodies are the served form's lattices (a box of cells on a ground plane) and
a strike is the served form's marble (banjo_amd/scenes.py), so a break here
is the same run a served trial is.
"""
from __future__ import annotations

import argparse
import json
import sys

from . import abi, scenes
from .live import FractureJob, LiveWork
from .serve import Trial
from .serve_wire import _vec3, serve_lines

# A break ends by the served trials' exit rule: calm, quiet and minimum
# substeps as a Trial has them.
EXIT_RULE = {k: Trial.__dataclass_fields__[k].default for k in ("calm_steps", "quiet_steps", "min_steps")}


def _result(r):
    return {"job": r.job, "body": r.name, "exit_reason": int(r.status.exit_reason),
            "substeps": int(r.status.total_steps), "broken_bonds": int(r.status.broken_bonds),
            "fragments": int(r.fragments), "abandoned": bool(r.abandoned), "cost_ms": float(r.cost_ms)}


class LiveWire:
    """The request dispatch of the wire over one LiveWork."""

    def __init__(self, work: LiveWork):
        self.work = work
        self.bodies = {}       # name -> (edges, seed)
        self.last = {}         # body -> its last finished result
        self.jobs = {}         # job id -> body

    def opening(self):
        return {"ok": True, "engine": "banjo_amd", "schema": "banjo.live.v1",
                "workers": self.work.workers, "dt_s": scenes.SERVE_DT,
                "max_substeps": scenes.SERVE_SUBSTEPS}

    def _collect(self, results):
        for r in results:
            self.last[r.name] = r
        return [_result(r) for r in results]

    def handle(self, cmd: dict) -> dict:
        if not isinstance(cmd, dict) or "op" not in cmd:
            raise ValueError('a request is an object with an "op"')
        op = cmd["op"]
        if op == "body":
            name = str(cmd["name"])
            if name in self.bodies:
                raise ValueError(f"body {name!r} exists")
            edges = [int(e) for e in cmd["edges"]]
            if len(edges) != 3 or min(edges) < 2 or max(edges) > 512:
                raise ValueError("edges are three cell counts in [2, 512]")
            seed = int(cmd.get("seed", 971))
            sc = scenes.served_lattice(edges, seed=seed, name=name)
            self.bodies[name] = (edges, seed)
            return {"ok": True, "body": name, "nodes": sc.node_count, "bonds": sc.bond_count}
        if op == "fracture":
            name = str(cmd["name"])
            if name not in self.bodies:
                raise ValueError(f"no body {name!r}")
            edges, seed = self.bodies[name]
            sc = scenes.served_lattice(edges, seed=seed, name=name)
            if "sphere" in cmd:
                b = cmd["sphere"]
                sp = scenes.sphere_state(center=_vec3(b["center_m"], "center_m"),
                                         velocity=_vec3(b.get("velocity_m_s", (0, 0, 0)), "velocity_m_s"),
                                         radius=float(b.get("radius_m", scenes.MARBLE_RADIUS)),
                                         mass=float(b.get("mass_kg", scenes.MARBLE_MASS)))
                if not (sp.radius > 0 and sp.mass > 0):
                    raise ValueError("the sphere needs a positive radius and mass")
            else:
                sp = scenes.served_marble(edges, float(cmd.get("speed_m_s", 1.0)), float(cmd.get("tilt_rad", 0.0)),
                                          float(cmd.get("azimuth_rad", 0.0)))
            sc.sphere = sp
            steps = int(cmd.get("max_steps", scenes.SERVE_SUBSTEPS))
            if not 0 < steps <= scenes.SERVE_SUBSTEPS:
                raise ValueError(f"max_steps is in [1, {scenes.SERVE_SUBSTEPS}]")
            k = self.work.submit(FractureJob(name=name, scene=sc, substeps=steps, control_kw=dict(EXIT_RULE)))
            self.jobs[k] = name
            if cmd.get("wait", True):
                r = self.work.wait(k)
                return {"ok": True, "fracture": name, "job": k, "result": self._collect([r])[0]}
            return {"ok": True, "fracture": name, "job": k, "pending": True}
        if op == "step":
            n = int(cmd.get("n", 1))
            if not 0 < n <= 100000:
                raise ValueError("n is in [1, 100000]")
            return {"ok": True, "running": self.work.running(), "fractures": self._collect(self.work.poll())}
        if op == "abandon":
            k = int(cmd["job"])
            if k not in self.jobs:
                raise ValueError(f"no job {k}")
            try:
                self.work.abandon(k)
            except KeyError:
                return {"ok": True, "job": k, "finished": True}
            return {"ok": True, "job": k, "abandoning": True}
        if op == "poses":
            cap = int(cmd.get("max_pieces", 64))
            pieces = []
            for name, r in sorted(self.last.items()):
                for f, m in enumerate(r.mass[:cap]):
                    pieces.append({"body": name, "job": r.job, "piece": f, "nodes": int(m["node_count"]),
                                   "mass_kg": float(m["mass_kg"]),
                                   "center_of_mass_m": [float(x) for x in m["center_of_mass_world_m"]],
                                   "velocity_m_s": [float(x) for x in m["linear_velocity_m_s"]],
                                   "angular_velocity_rad_s": [float(x) for x in m["angular_velocity_rad_s"]]})
            return {"ok": True, "pieces": pieces}
        if op == "quit":
            return {"ok": True, "bye": True}
        raise ValueError(f"unknown op {op!r}")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--workers", type=int, default=2)
    args = ap.parse_args(argv)
    from .lattice import product
    hip = product()
    if not hip.lib.blx_available():
        print(json.dumps({"ok": False, "error": "no gfx950 device visible to the HIP lattice library"}), flush=True)
        return 1
    work = LiveWork(hip, workers=args.workers, precision=abi.PRECISION_FLOAT)
    try:
        serve_lines(LiveWire(work), sys.stdin, sys.stdout)
    finally:
        work.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
