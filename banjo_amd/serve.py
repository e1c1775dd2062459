"""Continuous-batching server of lattice trials on one MI355X.

Trials arrive over time; each is a lattice already resident on the device as
a template (uploaded once, never run) plus its own marble and run control.
An arriving trial is admitted as soon as a batch slot is free and the device
bytes of the admitted trials stay within the HBM budget: admission clones the
template into the slot on the device and arms it. The batch advances every
running trial together in chunks of substeps; a trial that has stopped (its
exit rule or its 3 s budget) is answered with its fragment count and whether
the marble has settled, and its slot is freed for the next arrival. A trial's
results do not depend on what else is in the batch: each slot evolves exactly
as the same lattice would alone.
"""
from __future__ import annotations

import collections
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .lattice import Batch, LatticeLibrary, control
from .runner import prepare

SETTLED_SPEED = 0.01  # m/s: the marble is at rest (TileImpactScene.hpp rest_speed_m_s)


@dataclass
class Trial:
    id: int
    arrival: float          # seconds after the stream starts
    template: int
    sphere: abi.Sphere
    max_steps: int
    calm_steps: int = 480
    quiet_steps: int = 480
    min_steps: int = 480


@dataclass
class Answer:
    id: int
    arrival: float
    admitted: float
    finished: float
    fragments: int
    settled: bool
    exit_reason: int
    substeps: int
    broken_bonds: int
    slot: int
    state: object = field(default=None, repr=False)
    labels: np.ndarray | None = field(default=None, repr=False)

    @property
    def latency(self):
        return self.finished - self.arrival


class TrialServer:
    def __init__(self, lib: LatticeLibrary, templates, slots, hbm_budget_bytes=200 << 30, chunk=48,
                 precision=abi.PRECISION_FLOAT):
        self.lib = lib
        self.templates = templates
        self.slots = int(slots)
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk)
        T = len(templates)
        self.batch = Batch(lib, precision, T + self.slots)
        self.prepared = []
        for k, sc in enumerate(templates):
            sched, state = prepare(lib, sc)
            self.batch.upload(k, sched, state, sc.settings, sc.sphere)
            self.prepared.append((sched, state))
        self.footprint = [self.batch.slot_bytes(k) for k in range(T)]

    def close(self):
        self.batch.close()

    def control_of(self, t: Trial):
        return control(t.max_steps, calm_steps=t.calm_steps, quiet_steps=t.quiet_steps, min_steps=t.min_steps)

    def serve(self, trials, keep_state=False, clock=time.perf_counter, idle_sleep=2e-4):
        """Serves `trials` (sorted by arrival, arrival in seconds from now) and
        returns their answers in completion order."""
        T = len(self.templates)
        pending = collections.deque(sorted(trials, key=lambda t: t.arrival))
        free = list(range(T + self.slots - 1, T - 1, -1))
        active = {}
        used = 0
        answers = []
        t0 = clock()
        peak_active, peak_bytes = 0, 0
        self.seconds = collections.Counter()
        while pending or active:
            now = clock() - t0
            ta = clock()
            while pending and pending[0].arrival <= now and free:
                t = pending[0]
                need = self.footprint[t.template]
                if used + need > self.budget:
                    break
                pending.popleft()
                slot = free.pop()
                self.batch.clone(slot, t.template, t.sphere)
                self.batch.arm(slot, self.control_of(t))
                active[slot] = (t, clock() - t0, need)
                used += need
            self.seconds["admit"] += clock() - ta
            peak_active = max(peak_active, len(active))
            peak_bytes = max(peak_bytes, used)
            if not active:
                if pending:
                    wait = pending[0].arrival - (clock() - t0)
                    if wait > 0:
                        time.sleep(min(wait, idle_sleep))
                continue
            ta = clock()
            self.batch.advance(self.chunk)
            tb = clock()
            self.seconds["advance"] += tb - ta
            for slot in list(active):
                st = self.batch.status(slot)
                if st.exit_reason == 0:
                    continue
                t, admitted, need = active.pop(slot)
                labels, fragments = self.batch.fragment_labels(slot)
                sp = self.batch.sphere(slot)
                speed = math.sqrt(sum(x * x for x in sp.velocity))
                state = None
                if keep_state:
                    state = self.prepared[t.template][1].copy()
                    self.batch.download(slot, state)
                answers.append(Answer(id=t.id, arrival=t.arrival, admitted=admitted, finished=clock() - t0,
                                      fragments=fragments, settled=speed < SETTLED_SPEED,
                                      exit_reason=st.exit_reason, substeps=st.total_steps,
                                      broken_bonds=st.broken_bonds, slot=slot, state=state,
                                      labels=labels if keep_state else None))
                self.batch.release(slot)
                free.append(slot)
                used -= need
            self.seconds["retire"] += clock() - tb
        self.peak_active, self.peak_bytes = peak_active, peak_bytes
        return answers


def latency_summary(answers):
    lat = np.array([a.latency for a in answers])
    span = max(a.finished for a in answers) - min(a.arrival for a in answers)
    return {"trials": len(answers), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()), "trials_per_s": len(answers) / span,
            "mean_substeps": float(np.mean([a.substeps for a in answers])),
            "fragmented": int(sum(a.fragments > 1 for a in answers)),
            "settled": int(sum(a.settled for a in answers))}
