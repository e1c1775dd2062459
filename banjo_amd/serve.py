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


class TrialSession:
    """The server's state between calls: resident templates, the trials
    waiting for a slot, the trials on the device, and the HBM they hold.

    Templates are uploaded into the first `max_templates` batch slots (any
    time, up to that many); trials go into the remaining `slots`. `submit`
    queues a trial; `pump` admits queued trials in arrival order while a slot
    is free and the admitted bytes stay within the budget, advances every
    running trial one chunk, and answers and frees those that stopped.
    """

    def __init__(self, lib: LatticeLibrary, slots, max_templates, hbm_budget_bytes=200 << 30, chunk=48,
                 precision=abi.PRECISION_FLOAT, clock=time.perf_counter):
        self.lib = lib
        self.slots = int(slots)
        self.max_templates = int(max_templates)
        self.budget = int(hbm_budget_bytes)
        self.chunk = int(chunk)
        self.clock = clock
        self.batch = Batch(lib, precision, self.max_templates + self.slots)
        self.templates, self.prepared, self.footprint = [], [], []
        self.free = list(range(self.max_templates + self.slots - 1, self.max_templates - 1, -1))
        self.pending = collections.deque()
        self.active = {}
        self.used = 0
        self.t0 = clock()
        self.peak_active, self.peak_bytes = 0, 0
        self.seconds = collections.Counter()

    def close(self):
        self.batch.close()

    def now(self):
        return self.clock() - self.t0

    def add_template(self, scene) -> int:
        """Uploads a lattice as a template (never run; trials clone it)."""
        k = len(self.templates)
        if k >= self.max_templates:
            raise ValueError(f"at most {self.max_templates} templates")
        sched, state = prepare(self.lib, scene)
        self.batch.upload(k, sched, state, scene.settings, scene.sphere)
        need = self.batch.slot_bytes(k)
        if need > self.budget:
            # A trial of it could never be admitted: refuse the template
            # rather than queue trials that would wait for ever.
            self.batch.release(k)
            raise ValueError(f"template needs {need} bytes of HBM, above the {self.budget}-byte budget")
        self.templates.append(scene)
        self.prepared.append((sched, state))
        self.footprint.append(self.batch.slot_bytes(k))
        return k

    def submit(self, t: Trial):
        if not 0 <= t.template < len(self.templates):
            raise ValueError(f"no template {t.template}")
        if self.footprint[t.template] > self.budget:
            raise ValueError(f"template {t.template} does not fit the HBM budget")
        self.pending.append(t)

    @property
    def busy(self):
        return bool(self.pending or self.active)

    def control_of(self, t: Trial):
        return control(t.max_steps, calm_steps=t.calm_steps, quiet_steps=t.quiet_steps, min_steps=t.min_steps)

    def admit(self):
        ta = self.clock()
        while self.pending and self.free:
            t = self.pending[0]
            need = self.footprint[t.template]
            if self.used + need > self.budget:
                break
            self.pending.popleft()
            slot = self.free.pop()
            self.batch.clone(slot, t.template, t.sphere)
            self.batch.arm(slot, self.control_of(t))
            self.active[slot] = (t, self.now(), need)
            self.used += need
        self.seconds["admit"] += self.clock() - ta
        self.peak_active = max(self.peak_active, len(self.active))
        self.peak_bytes = max(self.peak_bytes, self.used)

    def pump(self, keep_state=False):
        """Admits, advances one chunk and returns the answers of the trials
        that stopped (in slot order)."""
        self.admit()
        if not self.active:
            return []
        ta = self.clock()
        self.batch.advance(self.chunk)
        tb = self.clock()
        self.seconds["advance"] += tb - ta
        answers = []
        # One call for every slot's exit reason; the full status only for those that stopped.
        reasons = self.batch.exit_reasons(self.max_templates + self.slots)
        for slot in [k for k in self.active if reasons[k] != 0]:
            st = self.batch.status(slot)
            t, admitted, need = self.active.pop(slot)
            labels, fragments = self.batch.fragment_labels(slot)
            sp = self.batch.sphere(slot)
            speed = math.sqrt(sum(x * x for x in sp.velocity))
            state = None
            if keep_state:
                state = self.prepared[t.template][1].copy()
                self.batch.download(slot, state)
            answers.append(Answer(id=t.id, arrival=t.arrival, admitted=admitted, finished=self.now(),
                                  fragments=fragments, settled=speed < SETTLED_SPEED, exit_reason=st.exit_reason,
                                  substeps=st.total_steps, broken_bonds=st.broken_bonds, slot=slot, state=state,
                                  labels=labels if keep_state else None))
            self.batch.release(slot)
            self.free.append(slot)
            self.used -= need
        self.seconds["retire"] += self.clock() - tb
        return answers


class TrialServer:
    """A stream of trials over fixed templates, each trial admitted once its
    arrival time has passed (the offered-load form of TrialSession)."""

    def __init__(self, lib: LatticeLibrary, templates, slots, hbm_budget_bytes=200 << 30, chunk=48,
                 precision=abi.PRECISION_FLOAT):
        self.session = TrialSession(lib, slots, len(templates), hbm_budget_bytes, chunk, precision)
        for sc in templates:
            self.session.add_template(sc)
        self.templates = templates
        self.footprint = self.session.footprint

    def close(self):
        self.session.close()

    @property
    def budget(self):
        return self.session.budget

    @budget.setter
    def budget(self, value):
        self.session.budget = int(value)

    def serve(self, trials, keep_state=False, clock=time.perf_counter, idle_sleep=2e-4):
        """Serves `trials` (sorted by arrival, arrival in seconds from now) and
        returns their answers in completion order."""
        s = self.session
        s.clock = clock
        s.t0 = clock()
        s.peak_active, s.peak_bytes = 0, 0
        s.seconds = collections.Counter()
        arriving = collections.deque(sorted(trials, key=lambda t: t.arrival))
        answers = []
        while arriving or s.busy:
            now = s.now()
            while arriving and arriving[0].arrival <= now:
                s.submit(arriving.popleft())
            if not s.busy:
                wait = arriving[0].arrival - s.now()
                if wait > 0:
                    time.sleep(min(wait, idle_sleep))
                continue
            answers += s.pump(keep_state)
        self.peak_active, self.peak_bytes, self.seconds = s.peak_active, s.peak_bytes, s.seconds
        return answers


def latency_summary(answers):
    lat = np.array([a.latency for a in answers])
    span = max(a.finished for a in answers) - min(a.arrival for a in answers)
    return {"trials": len(answers), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "mean_s": float(lat.mean()), "max_s": float(lat.max()), "trials_per_s": len(answers) / span,
            "mean_substeps": float(np.mean([a.substeps for a in answers])),
            "fragmented": int(sum(a.fragments > 1 for a in answers)),
            "settled": int(sum(a.settled for a in answers))}
