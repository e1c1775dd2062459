"""Runs a Scene through any library that exports the lattice C ABI.

The sequence is the reference harness's (src/fastlattice/TileImpactScene.cpp:777-868):
buildLatticeSchedule -> buildLatticeState -> backend upload -> external loads
-> run -> download -> writeBackLatticeState (bond state back to asset order)
-> findConnectedComponents.
"""
from __future__ import annotations

import time

import numpy as np

from . import abi
from .lattice import LatticeLibrary, control as make_control


def prepare(lib: LatticeLibrary, scene, origin=(0.0, 0.0, 0.0)):
    """Schedule and state for a scene, with clamps applied to the state."""
    sched = lib.schedule_build(scene.asset, scene.slabs)
    state = lib.state_build(scene.asset, sched, origin)
    if scene.clamped is not None:
        idx = np.nonzero(scene.clamped)[0]
        state.inv_mass[idx] = 0.0
        v = state.v.reshape(-1, 3)
        v[idx] = 0.0
    return sched, state


def run_scene(lib: LatticeLibrary, scene, precision=abi.PRECISION_FLOAT, substeps=None, backend_args=(),
              labels="host", chunks=None, control_kw=None, prepared=None, keep_backend=False, tuning=None,
              autotune=None):
    """Advance `scene` and return the results in asset order.

    labels: "host" (library's host components), "device" (backend fragment
    labels), or None. chunks: list of substep counts for successive run()
    calls (tests the status carrying over between calls). tuning: launch
    tuning fields to set (blx_tuning); autotune: (warm, probe) substeps to
    run the autotuner with before the run (its report is res["autotune"]).
    """
    sched, state = prepared if prepared is not None else prepare(lib, scene)
    be = lib.backend(precision, sched, *backend_args)
    be.upload(state, scene.settings, scene.sphere)
    if scene.info.get("blade") is not None:
        be.set_blade(scene.info["blade"])
    if scene.forces is not None:
        be.set_external_forces(scene.forces, scene.force_substeps)
    if tuning:
        be.set_tuning(**tuning)
    report = be.autotune(*autotune) if autotune else None
    total = scene.substeps if substeps is None else substeps
    plan = chunks if chunks is not None else [total]
    t0 = time.perf_counter()
    st = None
    for n in plan:
        st = be.run(make_control(n, **(control_kw or {})))
    wall = time.perf_counter() - t0
    out = state.copy()
    sphere = be.download(out)
    order = sched.bond_order
    B = scene.bond_count
    alive = np.empty(B, np.uint8)
    alive[order] = out.alive
    damage = np.empty(B)
    damage[order] = out.damage
    mode = np.empty(B, np.uint8)
    mode[order] = out.failure_mode
    res = dict(u=out.u.reshape(-1, 3), v=out.v.reshape(-1, 3), u_prev=out.u_prev.reshape(-1, 3),
               alive=alive, damage=damage, mode=mode, status=st, sphere=sphere, wall=wall,
               first_failure=np.sort(order[be.first_failure_bonds()]), schedule=sched, state=out, autotune=report)
    if labels == "host":
        res["labels"], res["components"] = lib.components(scene.node_count, scene.asset.bond_a, scene.asset.bond_b,
                                                          alive)
    elif labels == "device":
        res["labels"], res["components"] = be.fragment_labels()
    if keep_backend:
        res["backend"] = be
    else:
        be.close()
    return res


def collect(lib: LatticeLibrary, scene, sched, state_out, sphere, status, first_failure_sched, labels):
    """Results of a finished run in asset order (the same fields as run_scene)."""
    order = sched.bond_order
    B = scene.bond_count
    alive = np.empty(B, np.uint8)
    alive[order] = state_out.alive
    damage = np.empty(B)
    damage[order] = state_out.damage
    mode = np.empty(B, np.uint8)
    mode[order] = state_out.failure_mode
    res = dict(u=state_out.u.reshape(-1, 3), v=state_out.v.reshape(-1, 3), u_prev=state_out.u_prev.reshape(-1, 3),
               alive=alive, damage=damage, mode=mode, status=status, sphere=sphere, schedule=sched, state=state_out)
    if first_failure_sched is not None:
        res["first_failure"] = np.sort(order[first_failure_sched])
    if labels is not None:
        res["labels"], res["components"] = labels
    return res


def run_batch(lib: LatticeLibrary, scenes_, precision=abi.PRECISION_FLOAT, substeps=None, control_kw=None,
              steps_per_launch=0, capacity=None, labels=True, tuning=None, autotune=None):
    """Advance several scenes together in one batch (they must share step
    settings) and return each one's results as run_scene would. tuning and
    autotune as for run_scene (the autotuner runs once every trial is armed)."""
    from .lattice import Batch
    cap = capacity or len(scenes_)
    bt = Batch(lib, precision, cap)
    prepared = []
    try:
        for k, sc in enumerate(scenes_):
            sched, state = prepare(lib, sc)
            bt.upload(k, sched, state, sc.settings, sc.sphere)
            if sc.info.get("blade") is not None:
                bt.set_blade(sc.info["blade"])
            if sc.forces is not None:
                bt.set_external_forces(k, sc.forces, sc.force_substeps)
            n = sc.substeps if substeps is None else (substeps[k] if hasattr(substeps, "__len__") else substeps)
            kw = control_kw[k] if isinstance(control_kw, (list, tuple)) else (control_kw or {})
            bt.arm(k, make_control(n, **kw))
            prepared.append((sched, state))
        if tuning:
            bt.set_tuning(**tuning)
        report = bt.autotune(*autotune) if autotune else None
        t0 = time.perf_counter()
        while bt.advance(1 << 62, steps_per_launch):
            pass
        wall = time.perf_counter() - t0
        out = []
        for k, sc in enumerate(scenes_):
            sched, state = prepared[k]
            st_out = state.copy()
            sphere = bt.download(k, st_out)
            res = collect(lib, sc, sched, st_out, sphere, bt.status(k), bt.first_failure_bonds(k),
                          bt.fragment_labels(k) if labels else None)
            res["wall"] = wall
            res["autotune"] = report
            out.append(res)
        return out
    finally:
        bt.close()
