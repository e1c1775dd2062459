"""Fracture jobs on the device while the caller goes on: the live world's
work step for the lattice lane.

The reference's live world prepares a fracture job on its own thread (the
island, its schedule and state, a backend and its RunControl), hands it to a
worker that runs ``backend->run(control)`` and ``download`` ("the only part
that takes any time, and the only part that touches nothing shared",
src/fastlattice/LiveWorld.cpp:12727-12753), and applies the result back on
its own thread (``applyPending``, :12755). ``RunControl::abandon`` cancels a
run between launches (src/fastlattice/FastLattice.hpp:122-127).

Here a job is one lattice island: a Scene, whether a served lattice struck by
a marble or a fragment made a lattice again by ``LatticeLibrary.fragment_lattice``
(the island build, src/fastlattice/Refracture.cpp:267-393). ``LiveWork.submit``
queues it and returns at once. A pool of worker threads each takes a job,
makes its backend (its own HIP streams, so jobs that run together overlap on
the device), uploads, runs, downloads, and takes the fragment labels and
mass properties; ``poll`` returns the jobs finished so far, ``wait`` blocks
for one, and ``abandon`` sets the byte the job's run polls between device
batches. The workers share nothing but the device, so each job's result is
what the same island gives run alone.

With a library that has no device (the oracle), the labels and the mass
properties come from its host functions instead; the job is otherwise the same.
"""
from __future__ import annotations

import collections
import ctypes as C
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .lattice import LatticeLibrary, control
from .runner import prepare

EXIT_ABANDONED = 7  # RunStatus::exit_reason of a run its abandon flag stopped


@dataclass
class FractureJob:
    """What a fracture hands the worker: the island and its run."""
    name: str
    scene: object
    substeps: int | None = None        # None: the scene's own
    steps_per_launch: int = 48         # substeps between abandon polls
    control_kw: dict = field(default_factory=dict)


@dataclass
class FractureResult:
    """What the worker hands back (the reference's Pending after work)."""
    job: int
    name: str
    status: abi.RunStatus
    state: object                      # the downloaded LatticeState
    labels: np.ndarray                 # canonical fragment label of every node
    fragments: int
    mass: np.ndarray                   # abi.FRAGMENT_MASS_DTYPE, one row per fragment, label order
    cost_ms: float
    abandoned: bool


def _fragment_lists(labels):
    """Node lists in label order, ascending nodes per fragment."""
    order = np.argsort(labels, kind="stable").astype(np.uint32)
    counts = np.bincount(labels, minlength=int(labels.max()) + 1 if len(labels) else 0)
    return np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32), order


class LiveWork:
    """Fracture jobs run by `workers` threads, each job on its own backend."""

    def __init__(self, lib: LatticeLibrary, workers=2, precision=abi.PRECISION_FLOAT, backend_args=None):
        self.lib = lib
        self.precision = precision
        self.device = lib.has("backend_fragment_labels")
        self.backend_args = tuple(backend_args) if backend_args is not None else ((0,) if self.device else ())
        self.workers = int(workers)
        self.pool = ThreadPoolExecutor(max_workers=self.workers, thread_name_prefix="fracture")
        self.lock = threading.Lock()
        self.futures = {}                      # job id -> future
        self.flags = {}                        # job id -> abandon byte
        self.done = collections.deque()        # finished job ids, in completion order
        self.next_id = 1

    def close(self):
        with self.lock:
            for flag in self.flags.values():
                flag[0] = 1
        self.pool.shutdown(wait=True)

    def submit(self, job: FractureJob) -> int:
        """Queues a job; returns its id at once."""
        with self.lock:
            k = self.next_id
            self.next_id += 1
            flag = (C.c_uint8 * 1)(0)
            self.flags[k] = flag
            fut = self.pool.submit(self._work, k, job, flag)
            self.futures[k] = fut
        fut.add_done_callback(lambda _f, k=k: self._finished(k))
        return k

    def _finished(self, k):
        with self.lock:
            self.done.append(k)

    def abandon(self, k):
        """RunControl::abandon: the job's run stops at its next device batch."""
        with self.lock:
            if k not in self.flags:
                raise KeyError(f"no job {k}")
            self.flags[k][0] = 1

    def running(self):
        with self.lock:
            return sum(1 for f in self.futures.values() if not f.done())

    def poll(self):
        """The jobs finished since the last poll, as results (raises a job's error)."""
        out = []
        while True:
            with self.lock:
                if not self.done:
                    break
                k = self.done.popleft()
                fut = self.futures.pop(k, None)  # None: a wait() took it already
                self.flags.pop(k, None)
            if fut is not None:
                out.append(fut.result())
        return out

    def wait(self, k, timeout=None) -> FractureResult:
        """Blocks until job k is done and returns its result."""
        with self.lock:
            fut = self.futures[k]
        res = fut.result(timeout=timeout)
        with self.lock:
            self.futures.pop(k, None)
            self.flags.pop(k, None)
        return res

    # The worker: prepare, run, download, fragments. Nothing here is shared
    # with another job but the library.
    def _work(self, k, job: FractureJob, flag) -> FractureResult:
        lib, sc = self.lib, job.scene
        t0 = time.perf_counter()
        sched, state = prepare(lib, sc)
        be = lib.backend(self.precision, sched, *self.backend_args)
        try:
            be.upload(state, sc.settings, sc.sphere)
            if sc.info.get("blade") is not None:
                be.set_blade(sc.info["blade"])
            if sc.forces is not None:
                be.set_external_forces(sc.forces, sc.force_substeps)
            n = sc.substeps if job.substeps is None else job.substeps
            ctl = control(n, steps_per_launch=job.steps_per_launch,
                          abandon=C.cast(flag, C.POINTER(C.c_uint8)), **job.control_kw)
            st = be.run(ctl)
            out = state.copy()
            be.download(out)
            if self.device:
                labels, count = be.fragment_labels()
                mass = be.fragment_mass_properties(sc.cell)
            else:
                alive = np.empty(sc.bond_count, np.uint8)
                alive[sched.bond_order] = out.alive
                labels, count = lib.components(sc.node_count, sc.asset.bond_a, sc.asset.bond_b, alive)
                pos = (out.origin[None, :] + out.x0.reshape(-1, 3)) + out.u.reshape(-1, 3)
                off, nodes = _fragment_lists(labels)
                mass = lib.fragment_mass(pos, out.v.reshape(-1, 3), out.mass, sc.cell, off, nodes)
        finally:
            be.close()
        return FractureResult(job=k, name=job.name, status=st, state=out, labels=labels, fragments=int(count),
                              mass=mass, cost_ms=1e3 * (time.perf_counter() - t0),
                              abandoned=st.exit_reason == EXIT_ABANDONED)
