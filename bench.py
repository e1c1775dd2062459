"""Benchmark of the MI355X lattice step: bond-solves/s on the 1M-node impact block.

    python bench.py [--gpus 1] [--steps 120] [--warmup 12]

Workload: a simple-cubic block of 128 (x) x 64 (y, vertical) x 128 (z) nodes,
26-neighbour bonds (13,338,492), per-bond compliance log-uniform in
[1e-7, 1e-5] and Weibull(k=6, 1.5%) break strain from seed 971, struck on its
top-face centre by a 10 kg, 0.2 m sphere at 8 m/s, on a ground plane. One step
is 1/240 s of simulated time: 10 substeps of dt = 1/2400 s; --steps 120 is the
whole 0.5 s run. Float arithmetic (the reference's fast displacement path).

A bond-solve is one bond carried through one whole substep; the rate is
bonds x substeps / seconds, the reference's bond_updates_per_s
(src/fastlattice/TileImpactScene.cpp:867-868). Inputs are resident on the
device before the timed region; the timed region is the backend's run() call
bracketed by device synchronisation.

Printed JSON carries, besides the rate:
  roofline      the dominant kernel's algorithmic bytes per launch over its
                mean HIP-event duration (a second, profiled pass) against
                MI355X's 8 TB/s HBM peak;
  step_roofline the whole substep at the algorithmic bytes per bond-solve;
  cpu_baseline  the reference's threaded CPU lane (oracle/_ref) on this host's
                cores over a bounded window of the same workload;
  batched       512 cantilevers of uneven sizes in one batch against the plain
                backend on a sample of the same trials (trials/s, speed-up);
  served        a Poisson stream of marble trials admitted into a running
                batch (trials/s, p50/p99 latency);
  cut           a 256 x 32 x 32 plastic plank cut by a kinematic blade, with
                fragment labels after every 12-substep step;
  contact       the cut with node-node contact on, against off, over a window;
  c1            the 16^3 cantilever of config 1, run whole.
The CPU baseline also carries, under "more", the reference's threaded and
serial CPU lanes on config 1, config 2 and the plank.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
SUBSTEPS_PER_STEP = 10
METRIC = "bond-solves/s @1M particles (%HBM roof); trials/s + p99, batched vs plain"


def algorithmic_bytes(N, B, stages, sphere=True, plastic=False, iterations=1):
    """Algorithmic HBM bytes per kernel launch at fp32: every array of the API data
    model read or written once per pass, node gathers counted once per node per pass.

    Per bond / per node, by pass:
      predict          -         / u r+w 24, v r 12, u_prev w 12 (+ x0 12 with a sphere)
      sweep (x iter)   37 (+4+4) / u r+w 24, inv_mass 4
      velocity         -         / u, u_prev r 24, v w 12, flags 2
      end strain       42        / u 12, rinv 36, unmeasured 12, strain w 24
      sample + break   88 (+16)  / u 12, strain 24, valid 1
    Summed: 167 + 223 N/B per bond-solve (the same count for one iteration).
    """
    per = {}
    per["k_predict"] = N * (48 + (12 if sphere else 0))
    sweep_bond = 37 + (4 if iterations > 1 else 0) + (4 if plastic else 0)
    per["k_solve"] = (B * sweep_bond + N * 28) / max(stages, 1)
    per["k_sweep_pair"] = 2 * per["k_solve"]  # two colour stages per launch
    per["k_sweep_triple"] = 3 * per["k_solve"]  # three colour stages per launch
    per["k_velocity"] = N * 38
    per["k_strain"] = B * 42 + N * 84
    per["k_bond_end"] = B * (88 + (16 if plastic else 0)) + N * 37
    step = B * (167 + (4 if plastic else 0) + 16 * plastic) + N * (223 + (12 if sphere else 0))
    return per, step


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "impact128_hbm_traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs of this workload (each scaled by its copy-probe calibration)."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    hits = [v for k, v in t.get("kernels", {}).items() if k == kernel or k.startswith(kernel + "<")]
    if not hits:
        return None
    v = max(hits, key=lambda x: x.get("dispatches", 0))
    return {"bytes": v["bytes"], "fetch_bytes": v.get("fetch_bytes"), "write_bytes": v.get("write_bytes"),
            "source": os.path.relpath(TRAFFIC_FILE, ROOT)}


def tuned(obj, args):
    """Applies the --tuning fields (blx_tuning) to a backend or batch."""
    if args.tuning:
        obj.set_tuning(**json.loads(args.tuning))
    return obj


def make_workload(args):
    from banjo_amd import scenes
    sc = scenes.impact(128, 64, 128, substeps=args.steps * SUBSTEPS_PER_STEP, blocks=args.blocks,
                       name="impact_128x64x128")
    return sc


def gpu_leg(args, sc):
    from banjo_amd import abi
    from banjo_amd.lattice import control, product
    from banjo_amd.runner import prepare

    hip = product()
    if not hip.lib.blx_available():
        raise SystemExit("bench.py: no gfx950 device visible to the HIP lattice library")
    sched, state = prepare(hip, sc)
    be = tuned(hip.backend(abi.PRECISION_FLOAT, sched, 0), args)
    substeps = args.steps * SUBSTEPS_PER_STEP
    if args.warmup:
        be.upload(state, sc.settings, sc.sphere)
        be.run(control(args.warmup * SUBSTEPS_PER_STEP))
    be.upload(state, sc.settings, sc.sphere)
    import ctypes as C
    t0 = time.perf_counter()
    st = be.run(control(substeps))
    seconds = time.perf_counter() - t0
    done = st.total_steps
    result = dict(seconds=seconds, substeps=done, broken=st.broken_bonds, kernel_seconds=st.kernel_seconds,
                  launches=st.launches, contacts=st.contact.impulse_contacts,
                  overflow=st.contact.candidate_overflow)
    # Profiled pass: per-launch HIP events on the backend's own stream.
    prof_substeps = min(substeps, args.profile_substeps)
    be.upload(state, sc.settings, sc.sphere)
    hip.check(hip.lib.blx_backend_set_profiling(be.h, C.c_int(1)), "set_profiling")
    be.run(control(prof_substeps))
    kernels = {}
    count = C.c_uint32()
    hip.check(hip.lib.blx_backend_kernel_stats(be.h, C.c_uint32(0), None, C.c_size_t(0), None, None,
                                                C.byref(count)), "kernel_stats")
    for k in range(count.value):
        name = C.create_string_buffer(64)
        secs = C.c_double()
        launches = C.c_uint64()
        hip.check(hip.lib.blx_backend_kernel_stats(be.h, C.c_uint32(k), name, C.c_size_t(64), C.byref(secs),
                                                    C.byref(launches), C.byref(count)), "kernel_stats")
        if launches.value:
            kernels[name.value.decode()] = dict(seconds=secs.value, launches=launches.value)
    be.close()
    ranges = sched.range_end.astype(int) - sched.range_begin.astype(int)
    cc = max(sched.color_count, 1)
    # Stages per iteration: interior colours plus even/odd boundary colours that are non-empty.
    r = ranges.reshape(sched.block_count, 2, cc)
    stages = int((r[:, 0, :].sum(0) > 0).sum() + (r[0::2, 1, :].sum(0) > 0).sum() + (r[1::2, 1, :].sum(0) > 0).sum())
    result["kernels"] = kernels
    result["stages"] = stages
    result["prof_substeps"] = prof_substeps
    return result


def batched_leg(args):
    """The batched form: many cantilevers of uneven sizes advanced together in
    one engine, against the plain per-trial backend on a random sample of the
    same trials (whose results must be identical, bit for bit)."""
    import numpy as np
    from banjo_amd import abi, scenes
    from banjo_amd.lattice import Batch, control, product
    from banjo_amd.runner import prepare

    hip = product()
    T, sub = args.batch_trials, args.batch_substeps
    edges = scenes.batch_edges(971, T)
    t0 = time.perf_counter()
    scs = [scenes.config3_trial(t, edges=edges[t], substeps=sub) for t in range(T)]
    prepared = [prepare(hip, sc) for sc in scs]
    t_build = time.perf_counter() - t0
    bt = tuned(Batch(hip, abi.PRECISION_FLOAT, T), args)
    t0 = time.perf_counter()
    for k, (sc, (sched, state)) in enumerate(zip(scs, prepared)):
        bt.upload(k, sched, state, sc.settings, sc.sphere)
        bt.set_external_forces(k, sc.forces, sc.force_substeps)
        bt.arm(k, control(sub))
    t_upload = time.perf_counter() - t0
    # The autotuner picks colour stages per launch, the LDS sweep and its
    # workgroup size, and substeps per graph from probes of this batch (and
    # puts every lattice back as it was).
    tune = None
    if args.batch_autotune:
        warm, probe = (int(x) for x in args.batch_autotune.split(","))
        t0 = time.perf_counter()
        rep = bt.autotune(warm, probe)
        tune = {"chosen": rep["chosen"], "candidates": len(rep["tried"]), "warm_substeps": warm,
                "probe_substeps": probe, "seconds": time.perf_counter() - t0}
    t0 = time.perf_counter()
    while bt.advance(1 << 62, args.batch_steps_per_launch):
        pass
    t_run = time.perf_counter() - t0
    bonds = sum(sc.bond_count for sc in scs)
    rng = np.random.default_rng(971)
    sample = sorted(rng.choice(T, size=min(args.plain_sample, T), replace=False).tolist())
    t_plain = 0.0
    identical = True
    for k in sample:
        sc, (sched, state) = scs[k], prepared[k]
        be = tuned(hip.backend(abi.PRECISION_FLOAT, sched, 0), args)
        be.upload(state, sc.settings, sc.sphere)
        be.set_external_forces(sc.forces, sc.force_substeps)
        t0 = time.perf_counter()
        be.run(control(sub))
        t_plain += time.perf_counter() - t0
        a, b = state.copy(), state.copy()
        be.download(a)
        bt.download(k, b)
        identical &= bool(np.array_equal(a.u, b.u) and np.array_equal(a.v, b.v) and np.array_equal(a.alive, b.alive))
        be.close()
    broken = sum(bt.status(k).broken_bonds for k in range(T))
    bt.close()
    plain_tps = len(sample) / t_plain
    return {"trials": T, "substeps_per_trial": sub, "nodes": sum(sc.node_count for sc in scs), "bonds": bonds,
            "broken_bonds": broken, "seconds": t_run, "trials_per_s": T / t_run,
            "bond_solves_per_s": bonds * sub / t_run, "plain_trials_per_s": plain_tps,
            "speedup_vs_plain": (T / t_run) / plain_tps, "plain_sample": len(sample),
            "identical_to_plain": identical, "host_build_s": t_build, "upload_s": t_upload, "autotune": tune,
            "workload": "cantilevers, edges log-uniform in [8, 48] per axis, 26-neighbour, Weibull(6, 1.5%) "
                        "break strain, log-uniform compliance, tip load U(1, 50) N, dt 1/1920 s"}


def served_leg(args):
    """The served form: Poisson arrivals of marble trials over lattices of the
    batched form's size distribution, admitted into a continuously running
    batch as slots and the HBM budget allow; p50/p99 latency from arrival to
    answer."""
    from banjo_amd import scenes
    from banjo_amd.lattice import product
    from banjo_amd.serve import Trial, TrialServer, latency_summary

    hip = product()
    pool = args.serve_templates
    edges = scenes.batch_edges(4242, pool)
    t0 = time.perf_counter()
    lats = [scenes.served_lattice(e, seed=4242 + k) for k, e in enumerate(edges)]
    server = TrialServer(hip, lats, slots=args.serve_slots, hbm_budget_bytes=args.serve_hbm_gib << 30,
                         chunk=args.serve_chunk)
    tuned(server.session.batch, args)
    t_setup = time.perf_counter() - t0
    stream = scenes.served_stream(971, args.serve_rate, args.serve_seconds, pool)
    trials = [Trial(id=i, arrival=a, template=k, sphere=scenes.served_marble(edges[k], v, tilt, az),
                    max_steps=scenes.SERVE_SUBSTEPS)
              for i, (a, k, v, tilt, az) in enumerate(stream)]
    try:
        answers = server.serve(trials)
    finally:
        server.close()
    s = latency_summary(answers)
    bond_solves = sum(lats[t.template].bond_count * a.substeps for a, t in
                      zip(sorted(answers, key=lambda a: a.id), trials))
    span = max(a.finished for a in answers) - min(a.arrival for a in answers)
    s.update({"offered_rate_per_s": args.serve_rate, "stream_seconds": args.serve_seconds,
              "bond_solves_per_s": bond_solves / span, "peak_concurrent_trials": server.peak_active,
              "peak_admitted_bytes": server.peak_bytes, "hbm_budget_bytes": args.serve_hbm_gib << 30,
              "slots": args.serve_slots, "chunk_substeps": args.serve_chunk, "lattice_pool": pool,
              "setup_s": t_setup, "host_seconds": dict(server.seconds),
              "workload": "marble 5 cm, 50 g, 0.5-3 m/s within 0.3 rad of vertical, onto lattices with edges "
                          "log-uniform in [8, 48] on a ground plane; up to 3 s (5760 substeps of 1/1920 s), "
                          "ending early by the calm/quiet rules (480 substeps)"})
    return s


def cpu_threads():
    """Host threads for the CPU baseline: every core this process may run on
    (its affinity), within the per-GPU CPU share the box sets in
    OMP_NUM_THREADS when it sets one."""
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    limit = int(share) if share.isdigit() and int(share) > 0 else avail
    return max(1, min(avail, limit)), avail


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def cpu_leg(args, sc):
    """The reference's threaded CPU lane over bounded windows of the same
    workload: the first substeps from t = 0 (the first samples every node),
    then a warm window that follows them, which is the reported value."""
    from banjo_amd import abi
    from banjo_amd.lattice import LatticeLibrary, control
    from banjo_amd.runner import prepare
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libbanjo_ref.so")
    orc_path = os.path.join(ROOT, "oracle", "build", "liboracle.so")
    threads, avail = cpu_threads()
    cold_n, warm_n = args.cpu_substeps, args.cpu_warm_substeps
    if os.path.exists(ref_path):
        lib, kind, extra, cores = LatticeLibrary(ref_path, "ref_", "ref"), "reference", (1, threads), threads
    elif os.path.exists(orc_path):
        lib, kind, extra, cores = LatticeLibrary(orc_path, "orc_", "oracle"), "port", (), 1
        cold_n, warm_n = max(1, cold_n // 5), max(1, warm_n // 10)
    else:
        return None
    sched, state = prepare(lib, sc)
    be = lib.backend(abi.PRECISION_FLOAT, sched, *extra)
    be.upload(state, sc.settings, sc.sphere)
    t0 = time.perf_counter()
    st = be.run(control(cold_n))
    cold_s = time.perf_counter() - t0
    cold_steps = st.total_steps
    t0 = time.perf_counter()
    st = be.run(control(warm_n))
    warm_s = time.perf_counter() - t0
    warm_steps = st.total_steps - cold_steps
    be.close()
    value = sc.bond_count * warm_steps / warm_s
    lane = "ParallelCpuLatticeBackend" if kind == "reference" else "serial oracle"
    out = {"value": value, "unit": "bond-solves/s", "cores": cores, "kind": kind,
           "sample": f"substeps {cold_steps}-{cold_steps + warm_steps} of the same 128x64x128 impact, float, {lane} "
                     f"with {cores} thread(s): {warm_s:.2f} s (substeps 0-{cold_steps} from t=0, which include "
                     f"the first substep's start sample of every node: "
                     f"{sc.bond_count * cold_steps / cold_s:.3g} bond-solves/s in {cold_s:.2f} s)",
           "cold_value": sc.bond_count * cold_steps / cold_s, "cold_substeps": cold_steps,
           "warm_substeps": warm_steps, "host_cpus": os.cpu_count(), "affinity_cpus": avail,
           "thread_rule": "every core of the process's affinity, within OMP_NUM_THREADS (the per-GPU CPU share)",
           "cpu_model": cpu_model()}
    if kind == "reference" and not args.cpu_quick:
        out["more"] = cpu_more(lib, sc, threads)
    return out


def cpu_more(lib, c2, threads):
    """The reference's CPU lane on the other workloads, threaded and serial:
    config 1 in full, config 2 and the plank over fixed windows."""
    from banjo_amd import abi, scenes
    from banjo_amd.lattice import control
    from banjo_amd.runner import prepare

    def rate(sc, kind, n, precision=abi.PRECISION_FLOAT):
        sched, state = prepare(lib, sc)
        be = lib.backend(precision, sched, kind, threads if kind == 1 else 1)
        be.upload(state, sc.settings, sc.sphere)
        if sc.forces is not None:
            be.set_external_forces(sc.forces, sc.force_substeps)
        t0 = time.perf_counter()
        st = be.run(control(n))
        secs = time.perf_counter() - t0
        be.close()
        return {"value": sc.bond_count * st.total_steps / secs, "substeps": st.total_steps, "seconds": secs,
                "threads": threads if kind == 1 else 1}

    c1 = scenes.config1()
    plank = scenes.config5(substeps=100)
    return {
        "c1_full_threaded": rate(c1, 1, c1.substeps),
        "c1_serial": rate(c1, 0, 128),
        "c2_serial": rate(c2, 0, 1),
        "plank_threaded": rate(plank, 1, 100),
        "plank_serial": rate(plank, 0, 3),
        "note": "config 1: 16^3 cantilever, 1,920 substeps (serial: first 128); config 2: the 128x64x128 impact "
                "(serial: first substep); plank: the 256x32x32 plastic plank on its ledges without the blade "
                "(the reference has none), first 100 substeps (serial: 3); float, bond-solves/s; the first "
                "substep after upload samples every node, so short serial windows read low",
    }


def precision_leg(args):
    """Float against double on the device (the double lane is the reference's
    FP64 answer): config 1 in full, the impact before and into the regime where
    its block's candidate list overflows, and the cut over its first substeps."""
    from banjo_amd import precision, scenes
    from banjo_amd.lattice import product

    hip = product()
    out = {"eps": precision.EPS,
           "note": "max |du|, |dv| of the float lane against the double lane; bonds alive in one lane and broken "
                   "in the other, and how many lie within eps (relative) of their threshold in the lane where "
                   "they broke; the reference's bar is max |du| < 1e-6 m (tests/fast_lattice_tests.cpp:435-447)"}
    out["c1_full"] = precision.report(hip, scenes.config1())
    out["impact_first_9"] = precision.report(hip, scenes.config2(substeps=9))
    out["impact_first_30"] = precision.report(hip, scenes.config2(substeps=30))
    out["cut_first_72"] = precision.report(hip, scenes.config5(substeps=72))
    return out


def copy_probe(nbytes, repeats):
    import ctypes as C
    from banjo_amd.lattice import product
    hip = product()
    secs = C.c_double()
    hip.check(hip.lib.blx_hbm_copy_probe(C.c_int(0), C.c_uint64(nbytes), C.c_uint32(repeats), C.byref(secs)),
              "hbm_copy_probe")
    return {"copy_bytes": nbytes, "read_bytes": nbytes, "write_bytes": nbytes, "seconds": secs.value,
            "gbps_read_plus_write": 2 * nbytes / secs.value / 1e9}


def contact_leg(args):
    """Node-node contact on the cut plank: the same plank and blade with
    contact on (radius half a cell, skin a quarter), so the crack faces and
    the blade's two halves do not pass through each other; the same window
    without contact beside it."""
    from banjo_amd import abi, scenes
    from banjo_amd.lattice import control, product
    from banjo_amd.runner import prepare

    hip = product()
    steps, per = args.contact_steps, 12
    out = {}
    for label, mode in (("without", 0), ("with", 2)):
        sc = scenes.config5(substeps=steps * per)
        if mode:
            scenes.with_node_contact(sc.settings, sc.cell, sc.node_count, mode)
        sched, state = prepare(hip, sc)
        be = tuned(hip.backend(abi.PRECISION_FLOAT, sched, 0), args)
        be.upload(state, sc.settings, sc.sphere)
        be.set_blade(sc.info["blade"])
        t0 = time.perf_counter()
        st = be.run(control(steps * per))
        secs = time.perf_counter() - t0
        be.close()
        nc = st.node_contact
        out[label] = {"seconds": secs, "bond_solves_per_s": sc.bond_count * st.total_steps / secs,
                      "broken_bonds": st.broken_bonds}
        if mode:
            out[label].update({"rebuilds": nc.rebuilds, "pairs_listed": nc.pairs_listed, "pair_tests": nc.pair_tests,
                               "contacts": nc.contacts, "pair_overflow": nc.pair_overflow,
                               "maximum_overlap_m": nc.maximum_overlap_m,
                               "maximum_position_correction_m": nc.maximum_position_correction_m})
    out["substeps"] = steps * per
    out["workload"] = ("the 256x32x32 plastic plank cut by the blade (as the cut form), first "
                       f"{steps * per} substeps, node contact on against off; no labels")
    return out


def c1_leg(args):
    """Config 1 on the GPU: the 16^3 cantilever, 1,920 substeps, run whole."""
    from banjo_amd import abi, scenes
    from banjo_amd.lattice import control, product
    from banjo_amd.runner import prepare

    hip = product()
    sc = scenes.config1()
    sched, state = prepare(hip, sc)
    be = tuned(hip.backend(abi.PRECISION_FLOAT, sched, 0), args)
    out = {}
    for label in ("warm", "timed"):
        be.upload(state, sc.settings, sc.sphere)
        be.set_external_forces(sc.forces, sc.force_substeps)
        t0 = time.perf_counter()
        st = be.run(control(sc.substeps))
        out[label] = (time.perf_counter() - t0, st)
    be.close()
    secs, st = out["timed"]
    return {"nodes": sc.node_count, "bonds": sc.bond_count, "substeps": st.total_steps, "seconds": secs,
            "bond_solves_per_s": sc.bond_count * st.total_steps / secs, "broken_bonds": st.broken_bonds,
            "workload": "16^3 cantilever clamped at x=0, 5 N tip load and gravity, c 1e-6, break strain 1%, "
                        "dt 1/1920 s, 1,920 substeps; a lattice this small is launch-bound (about 40 launches "
                        "per substep)"}


def slabs_leg(args):
    """The headline impact with its nodes in four z-slabs instead of one
    block (the reference's slab schedule, TileImpactScene.hpp:283 defaults to
    one): each slab lists its own 1,024 striker candidates, so the sphere meets
    up to four times as many nodes. Same window as the headline run."""
    from banjo_amd import abi, scenes
    from banjo_amd.lattice import control, product
    from banjo_amd.runner import prepare

    hip = product()
    substeps = args.steps * SUBSTEPS_PER_STEP
    out = {}
    for blocks in (1, 4):
        sc = scenes.impact(128, 64, 128, substeps=substeps, blocks=blocks, name=f"impact_128x64x128_{blocks}")
        sched, state = prepare(hip, sc)
        be = tuned(hip.backend(abi.PRECISION_FLOAT, sched, 0), args)
        be.upload(state, sc.settings, sc.sphere)
        t0 = time.perf_counter()
        st = be.run(control(substeps))
        secs = time.perf_counter() - t0
        be.close()
        out[f"blocks_{blocks}"] = {"seconds": secs, "bond_solves_per_s": sc.bond_count * st.total_steps / secs,
                                   "broken_bonds": st.broken_bonds,
                                   "striker_contacts": st.contact.impulse_contacts,
                                   "candidate_overflow": st.contact.candidate_overflow}
    out["substeps"] = substeps
    out["workload"] = "the headline impact from t = 0, one block against four z-slabs"
    return out


def cut_leg(args):
    """The cut form: a 256 x 32 x 32 plank on two ledges, plastic, cut at
    mid-span by a 1 mm kinematic blade moving down at 2 m/s. Every step is 12
    substeps followed by fragment labels computed on the device."""
    from banjo_amd import abi, scenes
    from banjo_amd.lattice import control, product
    from banjo_amd.runner import prepare

    hip = product()
    steps, per = args.cut_steps, 12
    t0 = time.perf_counter()
    sc = scenes.config5(substeps=steps * per)
    sched, state = prepare(hip, sc)
    t_build = time.perf_counter() - t0
    be = tuned(hip.backend(abi.PRECISION_FLOAT, sched, 0), args)
    be.upload(state, sc.settings, sc.sphere)
    be.set_blade(sc.info["blade"])
    t_sim = t_label = 0.0
    fragments = []
    st = None
    for _ in range(steps):
        t0 = time.perf_counter()
        st = be.run(control(per))
        t1 = time.perf_counter()
        _, count = be.fragment_labels()
        t_label += time.perf_counter() - t1
        t_sim += t1 - t0
        fragments.append(count)
    # The last step's labels against the host library's union-find over the same bonds.
    import numpy as np
    labels, _ = be.fragment_labels()
    out = state.copy()
    be.download(out)
    alive = np.empty(sc.bond_count, np.uint8)
    alive[sched.bond_order] = out.alive
    host, _ = hip.components(sc.node_count, sc.asset.bond_a, sc.asset.bond_b, alive)
    be.close()
    B = sc.bond_count
    total = t_sim + t_label
    return {"nodes": sc.node_count, "bonds": B, "steps": steps, "substeps": st.total_steps,
            "seconds": total, "simulate_s": t_sim, "label_s": t_label,
            "label_ms_per_step": t_label / steps * 1e3,
            "bond_solves_per_s": B * st.total_steps / total,
            "bond_solves_per_s_without_labels": B * st.total_steps / t_sim,
            "broken_bonds": st.broken_bonds, "broken_fraction": st.broken_bonds / B,
            "plastic_work_j": st.plastic_work_j, "fragments_final": fragments[-1],
            "fragments_max": max(fragments), "labels_match_host": bool(np.array_equal(labels, host)),
            "host_build_s": t_build,
            "workload": "plank 256x32x32, 26-neighbour, cell 1 cm, on two 16-cell ledges, gravity; plastic rest "
                        "length (yield 0.4%, hardening 0.1), Weibull(4, 2%) break strain; 1 mm blade at 2 m/s "
                        "down through mid-span; dt 1/2880 s, 12 substeps per step, device fragment labels after "
                        "every step"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=120)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--blocks", type=int, default=1)
    ap.add_argument("--profile-substeps", type=int, default=1200)
    ap.add_argument("--cpu-substeps", type=int, default=10)
    ap.add_argument("--cpu-warm-substeps", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-quick", action="store_true", help="only the config-2 window of the CPU baseline")
    ap.add_argument("--forms", default="plain,batched,served,cut,contact,c1,precision,slabs",
                    help="comma list of plain (the headline impact), batched (many uneven lattices), "
                         "served (a stream of trials), cut (a plank cut by a blade, labels every step), "
                         "contact (the cut with node-node contact on), c1 (the 16^3 cantilever in full), "
                         "precision (float against double on the device), slabs (the impact in one block "
                         "and in four z-slabs)")
    ap.add_argument("--cut-steps", type=int, default=600)
    ap.add_argument("--contact-steps", type=int, default=120)
    ap.add_argument("--batch-trials", type=int, default=512)
    ap.add_argument("--batch-substeps", type=int, default=8 * 240)
    ap.add_argument("--batch-steps-per-launch", type=int, default=0)
    ap.add_argument("--batch-autotune", default="8,24",
                    help="warm,probe substeps of the batched form's autotuner run before its timed region ('' for none)")
    ap.add_argument("--tuning", default="",
                    help="launch tuning fields (blx_tuning) as JSON, set on every backend and batch")
    ap.add_argument("--plain-sample", type=int, default=8)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-templates", type=int, default=64)
    ap.add_argument("--serve-slots", type=int, default=768)
    ap.add_argument("--serve-hbm-gib", type=int, default=200)
    ap.add_argument("--serve-chunk", type=int, default=48)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the HBM copy probe of this size (counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe_mib:
        print(json.dumps(copy_probe(args.copy_probe_mib << 20, 20)))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(the lattice step would shard by z-slabs across GPUs, which is designed but not built)",
              file=sys.stderr)
        raise SystemExit(2)
    forms = set(args.forms.split(","))
    if "plain" not in forms:
        out = {"metric": METRIC, "n_gpus": 1}
        if "batched" in forms:
            out["batched"] = batched_leg(args)
        if "served" in forms:
            out["served"] = served_leg(args)
        if "cut" in forms:
            out["cut"] = cut_leg(args)
        if "contact" in forms:
            out["contact"] = contact_leg(args)
        if "c1" in forms:
            out["c1"] = c1_leg(args)
        if "precision" in forms:
            out["precision"] = precision_leg(args)
        if "slabs" in forms:
            out["slabs"] = slabs_leg(args)
        print(json.dumps(out))
        return
    t_build = time.perf_counter()
    sc = make_workload(args)
    N, B = sc.node_count, sc.bond_count
    print(f"# workload {sc.name}: N={N} B={B} built in {time.perf_counter() - t_build:.1f} s", file=sys.stderr,
          flush=True)
    g = gpu_leg(args, sc)
    print(f"# gpu: {g['substeps']} substeps in {g['seconds']:.3f} s, broken {g['broken']}", file=sys.stderr, flush=True)
    rate = B * g["substeps"] / g["seconds"]
    per, step_bytes = algorithmic_bytes(N, B, g["stages"], sphere=True)
    kernels = g["kernels"]
    dominant = max(kernels, key=lambda k: kernels[k]["seconds"])
    avg = kernels[dominant]["seconds"] / kernels[dominant]["launches"]
    bytes_launch = per.get(dominant)
    roofline = None
    if bytes_launch is not None:
        achieved = bytes_launch / avg
        measured = measured_traffic(dominant)
        roofline = {"bound": "hbm", "kernel": dominant, "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
                    "unit": "GB/s", "frac": achieved / HBM_PEAK,
                    "traffic": measured["bytes"] if measured else None,
                    "bytes_per_launch": bytes_launch, "mean_launch_us": avg * 1e6}
        if measured:
            roofline["traffic_detail"] = measured
    per_bond_solve = step_bytes / B
    step_roofline = {"bytes_per_bond_solve": per_bond_solve, "achieved": per_bond_solve * rate / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": per_bond_solve * rate / HBM_PEAK}
    kernel_table = {k: {"mean_us": v["seconds"] / v["launches"] * 1e6, "launches": v["launches"],
                        "share": v["seconds"] / sum(x["seconds"] for x in kernels.values())}
                    for k, v in kernels.items()}
    probe = copy_probe(1 << 30, 10)
    cpu = None if args.no_cpu else cpu_leg(args, sc)
    out = {
        "metric": METRIC, "value": rate, "unit": "bond-solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": g["seconds"] / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "impact 128x64x128 (x, y vertical, z), 10 kg r=0.2 m sphere at 8 m/s, "
                               "dt 1/2400 s, 10 substeps per step", "nodes": N, "bonds": B,
                   "substeps": g["substeps"], "blocks": args.blocks, "colour_stages": g["stages"],
                   "broken_bonds": g["broken"], "striker_contacts": g["contacts"],
                   "candidate_overflow": g["overflow"]},
        "roofline": roofline, "step_roofline": step_roofline, "kernels": kernel_table,
        "copy_probe_gbps": probe["gbps_read_plus_write"],
        "cpu_baseline": cpu,
    }
    if "batched" in forms:
        out["batched"] = batched_leg(args)
    if "served" in forms:
        out["served"] = served_leg(args)
    if "cut" in forms:
        out["cut"] = cut_leg(args)
    if "contact" in forms:
        out["contact"] = contact_leg(args)
    if "c1" in forms:
        out["c1"] = c1_leg(args)
    if "precision" in forms:
        out["precision"] = precision_leg(args)
    if "slabs" in forms:
        out["slabs"] = slabs_leg(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
