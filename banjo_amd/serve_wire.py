"""Native lattice trials over JSON lines: the served form's wire.

The reference's engine process holds a world open and speaks JSON over stdin
and stdout, one object per line: an opening reply, then one reply per
request, ``{"ok":false,"error":...}`` for a request that failed, empty lines
skipped (tools/live_world_run.cpp:1-10, 2115-2125, 3558-3566). This process
speaks that transport for the served form: lattice templates resident on the
device, trials that arrive whenever the host sends them, continuous batching
(banjo_amd/serve.py TrialSession) and one answer per trial.

    python -m banjo_amd.serve_wire [--slots 768] [--max-templates 64] [--hbm-gib 200] [--chunk 48]

  in   {"op":"template","name":"small","edges":[16,12,20],"seed":971}
       {"op":"trial","id":"a1","template":"small","speed_m_s":2.0,"tilt_rad":0.1,"azimuth_rad":1.2}
       {"op":"trial","id":"a2","template":"small","max_steps":960,
        "sphere":{"center_m":[0,0.2,0],"velocity_m_s":[0,-2,0],"radius_m":0.05,"mass_kg":0.05}}
       {"op":"step","n":4}      advance every running trial n chunks (admitting as slots free)
       {"op":"drain"}           advance until every trial sent so far is answered
       {"op":"stats"}           latency percentiles and throughput of the answers so far
       {"op":"quit"}
  out  {"ok":true,"engine":"banjo_amd","slots":768,...}           on opening
       {"ok":true,"template":"small","index":0,"nodes":3840,"bonds":...,"bytes":...}
       {"ok":true,"queued":"a1","pending":1,"running":0}
       {"ok":true,"t_s":1.25,"running":3,"pending":0,"answers":[{"id":"a1","latency_s":0.41,
         "queued_s":0.0,"fragments":1,"settled":true,"exit_reason":1,"substeps":960,"broken_bonds":0}]}

A trial's answer does not depend on what else shares the batch: each slot
evolves exactly as the same lattice alone on the plain backend would.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np

from . import abi, scenes
from .serve import Trial, TrialSession, latency_summary


def _vec3(x, what):
    v = [float(c) for c in x]
    if len(v) != 3 or not all(np.isfinite(v)):
        raise ValueError(f"{what} must be three finite numbers")
    return v


class Wire:
    """One open session and the request dispatch of the JSON-lines wire."""

    def __init__(self, session: TrialSession):
        self.session = session
        self.names = {}          # template name -> index
        self.edges = {}          # template index -> lattice edges
        self.answers = []        # every answer so far (for stats)

    def opening(self):
        s = self.session
        return {"ok": True, "engine": "banjo_amd", "schema": "banjo.trials.v1", "slots": s.slots,
                "max_templates": s.max_templates, "hbm_budget_bytes": s.budget, "chunk_substeps": s.chunk,
                "dt_s": scenes.SERVE_DT, "max_substeps": scenes.SERVE_SUBSTEPS}

    def _answer(self, a):
        return {"id": a.id, "latency_s": a.latency, "queued_s": a.admitted - a.arrival, "fragments": int(a.fragments),
                "settled": bool(a.settled), "exit_reason": int(a.exit_reason), "substeps": int(a.substeps),
                "broken_bonds": int(a.broken_bonds)}

    def _progress(self, answers):
        self.answers += answers
        s = self.session
        return {"ok": True, "t_s": s.now(), "running": len(s.active), "pending": len(s.pending),
                "answers": [self._answer(a) for a in answers]}

    def handle(self, cmd: dict) -> dict:
        if not isinstance(cmd, dict) or "op" not in cmd:
            raise ValueError('a request is an object with an "op"')
        op = cmd["op"]
        s = self.session
        if op == "template":
            name = str(cmd["name"])
            if name in self.names:
                raise ValueError(f"template {name!r} exists")
            edges = [int(e) for e in cmd["edges"]]
            if len(edges) != 3 or min(edges) < 2 or max(edges) > 512:
                raise ValueError("edges are three cell counts in [2, 512]")
            sc = scenes.served_lattice(edges, seed=int(cmd.get("seed", 971)), name=name)
            k = s.add_template(sc)
            self.names[name] = k
            self.edges[k] = edges
            return {"ok": True, "template": name, "index": k, "nodes": sc.node_count, "bonds": sc.bond_count,
                    "bytes": int(s.footprint[k])}
        if op == "trial":
            name = str(cmd["template"])
            if name not in self.names:
                raise ValueError(f"no template {name!r}")
            k = self.names[name]
            if "sphere" in cmd:
                b = cmd["sphere"]
                sp = scenes.sphere_state(center=_vec3(b["center_m"], "center_m"),
                                         velocity=_vec3(b.get("velocity_m_s", (0, 0, 0)), "velocity_m_s"),
                                         radius=float(b.get("radius_m", scenes.MARBLE_RADIUS)),
                                         mass=float(b.get("mass_kg", scenes.MARBLE_MASS)))
                if not (sp.radius > 0 and sp.mass > 0):
                    raise ValueError("the sphere needs a positive radius and mass")
            else:
                sp = scenes.served_marble(self.edges[k], float(cmd.get("speed_m_s", 1.0)),
                                          float(cmd.get("tilt_rad", 0.0)), float(cmd.get("azimuth_rad", 0.0)))
            steps = int(cmd.get("max_steps", scenes.SERVE_SUBSTEPS))
            if not 0 < steps <= scenes.SERVE_SUBSTEPS:
                raise ValueError(f"max_steps is in [1, {scenes.SERVE_SUBSTEPS}]")
            s.submit(Trial(id=cmd.get("id", len(self.answers) + len(s.pending) + len(s.active)), arrival=s.now(),
                           template=k, sphere=sp, max_steps=steps))
            return {"ok": True, "queued": cmd.get("id"), "pending": len(s.pending), "running": len(s.active)}
        if op == "step":
            n = int(cmd.get("n", 1))
            if not 0 < n <= 100000:
                raise ValueError("n is in [1, 100000]")
            answers = []
            for _ in range(n):
                if not s.busy:
                    break
                answers += s.pump()
            return self._progress(answers)
        if op == "drain":
            answers = []
            while s.busy:
                answers += s.pump()
            return self._progress(answers)
        if op == "stats":
            out = {"ok": True, "answered": len(self.answers), "running": len(s.active), "pending": len(s.pending),
                   "peak_running": s.peak_active, "peak_bytes": int(s.peak_bytes)}
            if self.answers:
                out.update(latency_summary(self.answers))
            return out
        if op == "quit":
            return {"ok": True, "bye": True}
        raise ValueError(f"unknown op {op!r}")


def serve_lines(wire: Wire, lines, out):
    """The loop of the wire: one reply line per request line."""
    out.write(json.dumps(wire.opening()) + "\n")
    out.flush()
    for line in lines:
        line = line.strip()
        if not line:
            continue
        try:
            cmd = json.loads(line)
            reply = wire.handle(cmd)
        except Exception as e:  # a failed request is answered, and the session goes on
            reply = {"ok": False, "error": f"{type(e).__name__}: {e}"}
            cmd = None
        out.write(json.dumps(reply) + "\n")
        out.flush()
        if isinstance(cmd, dict) and cmd.get("op") == "quit":
            break


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--slots", type=int, default=768)
    ap.add_argument("--max-templates", type=int, default=64)
    ap.add_argument("--hbm-gib", type=int, default=200)
    ap.add_argument("--chunk", type=int, default=48)
    args = ap.parse_args(argv)
    from .lattice import product
    hip = product()
    if not hip.lib.blx_available():
        print(json.dumps({"ok": False, "error": "no gfx950 device visible to the HIP lattice library"}), flush=True)
        return 1
    session = TrialSession(hip, args.slots, args.max_templates, args.hbm_gib << 30, args.chunk, abi.PRECISION_FLOAT)
    try:
        serve_lines(Wire(session), sys.stdin, sys.stdout)
    finally:
        session.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
