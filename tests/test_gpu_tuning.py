"""Launch tuning and the autotuner, on the device.

Every blx_tuning setting must give the reference's bits: colour stages one or
two per launch, the whole sweep in one workgroup with the nodes in LDS (the
reference GPU lane's single-block shared-memory mode, FastLattice.hpp:229-231)
at each workgroup size, and substep graphs. Each setting is compared with the
oracle (the reference's CPU lane restated) on impacts, a plastic run, direct
arithmetic, the config-1 cantilever with its tip loads, in float and double,
over lattices that fit in LDS and lattices that do not; a batch mixes both.
RunStatus::shared_memory_* say which lattices swept in LDS. The autotuner
tries candidates from the lattice's current state and puts it back: a run
after autotuning (at the start or in the middle) equals one without.
"""
import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.lattice import control
from banjo_amd.runner import prepare, run_batch, run_scene
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu
F, D = abi.PRECISION_FLOAT, abi.PRECISION_DOUBLE
STAGES = {"colours_per_launch": 1, "lds_sweep": 0, "lds_threads": 1024, "graph_substeps": 0, "streams": 0}
VARIANTS = {
    "stages": {},
    "pairs": {"colours_per_launch": 2},
    "lds256": {"lds_sweep": 1, "lds_threads": 256},
    "lds1024": {"lds_sweep": 1, "lds_threads": 1024},
    "lds512_graphs8": {"lds_sweep": 1, "lds_threads": 512, "graph_substeps": 8},
    "pairs_graphs32": {"colours_per_launch": 2, "graph_substeps": 32},
    "split": {"streams": 1},
    "pairs_split_graphs8": {"colours_per_launch": 2, "streams": 1, "graph_substeps": 8},
    "one_stream": {"streams": 2},
}


def scene(case):
    if case == "impact":
        return scenes.small_impact(nx=16, ny=8, nz=16, substeps=120)
    if case == "impact_3slabs":
        return scenes.small_impact(nx=16, ny=10, nz=18, substeps=100, blocks=3)
    if case == "plastic":
        return scenes.plastic_impact(substeps=80)
    if case == "direct":
        sc = scenes.small_impact(nx=16, ny=8, nz=16, substeps=80)
        sc.settings.direct_arithmetic = 1
        return sc
    if case == "config1":
        return scenes.config1(substeps=160)
    if case == "large":
        return scenes.small_impact(nx=28, ny=12, nz=28, substeps=60)
    if case == "prestressed":
        # Rest lengths 1.6 times the node spacing: the paired sweep's tiles
        # (16 x 8 x 8 cells of the shortest rest length) then hold 2,880
        # nodes each, more than a workgroup, and its plan splits them in halves.
        sc = scenes.small_impact(nx=24, ny=10, nz=24, substeps=40)
        sc.asset.rest_length = sc.asset.rest_length * 1.6
        return sc
    raise KeyError(case)


def lds_bytes(gpu):
    text = gpu.description()
    return int(text.split(" KiB LDS per workgroup")[0].rsplit(", ", 1)[1]) << 10


@pytest.fixture(scope="module")
def oracle_runs(oracle):
    return {}


@pytest.mark.parametrize("precision", [F, D], ids=["float", "double"])
@pytest.mark.parametrize("case", ["impact", "impact_3slabs", "plastic", "direct", "config1", "large", "prestressed"])
def test_every_tuning_equals_the_oracle(gpu, oracle, oracle_runs, case, precision):
    sc = scene(case)
    key = (case, precision)
    if key not in oracle_runs:
        oracle_runs[key] = run_scene(oracle, sc, precision)
    want = oracle_runs[key]
    limit = lds_bytes(gpu)
    width = 16 if precision == F else 32
    fits = sc.node_count * width <= limit
    for name, t in VARIANTS.items():
        print(f"{case} {'float' if precision == F else 'double'} N={sc.node_count} (LDS {limit >> 10} KiB, "
              f"fits={fits}): {name}", flush=True)
        got = run_scene(gpu, sc, precision, tuning=dict(STAGES, **t), chunks=[37, 3, sc.substeps - 40])
        assert_same(got, want, ledger_rtol=max(1e-12, sc.node_count * 2.0 ** -52))
        in_lds = bool(t.get("lds_sweep")) and fits
        assert got["status"].shared_memory_positions == in_lds, name
        assert got["status"].shared_memory_bytes == (sc.node_count * width if in_lds else 0), name
    assert want["status"].broken_bonds > 0 or case == "config1"


def test_config1_sweeps_in_lds_by_default(gpu):
    sc = scenes.config1(substeps=8)
    res = run_scene(gpu, sc, F)
    assert res["status"].shared_memory_positions == (sc.node_count * 16 <= lds_bytes(gpu))


def test_batch_of_mixed_sizes_every_tuning_equal(gpu):
    # 40 x 8 x 40 (12,800 nodes) does not fit one workgroup's LDS in float:
    # with the LDS sweep on, it sweeps by stages while the rest sweep in LDS
    # on a stream of their own.
    trials = [scenes.small_impact(nx=n, ny=8, nz=n, substeps=48) for n in (8, 12, 16, 26, 30, 40)]
    limit = lds_bytes(gpu)
    base = run_batch(gpu, trials, F, tuning=STAGES, steps_per_launch=16)
    for name, t in VARIANTS.items():
        print(f"batch of {[sc.node_count for sc in trials]}: {name}", flush=True)
        got = run_batch(gpu, trials, F, tuning=dict(STAGES, **t), steps_per_launch=16)
        for sc, a, b in zip(trials, got, base):
            assert_same(a, b)
            in_lds = bool(t.get("lds_sweep")) and sc.node_count * 16 <= limit
            assert a["status"].shared_memory_positions == in_lds
    assert any(r["status"].broken_bonds > 0 for r in base)


def test_autotune_changes_no_result(gpu):
    sc = scenes.small_impact(nx=16, ny=8, nz=16, substeps=120)
    plain = run_scene(gpu, sc, F)
    tuned = run_scene(gpu, sc, F, autotune=(8, 16))
    assert_same(tuned, plain)
    rep = tuned["autotune"]
    assert len(rep["tried"]) >= 6
    chosen = rep["chosen"]
    assert any({k: v for k, v in t.items() if k != "us_per_substep"} == chosen for t in rep["tried"])
    best = min(rep["tried"], key=lambda t: t["us_per_substep"])
    assert {k: v for k, v in best.items() if k != "us_per_substep"} == chosen
    print("autotune:", chosen, [(t["colours_per_launch"], t["lds_sweep"], t["lds_threads"], t["graph_substeps"],
                                 round(t["us_per_substep"], 1)) for t in rep["tried"]], flush=True)


def test_autotune_in_the_middle_of_a_run_changes_no_result(gpu):
    sc = scenes.small_impact(nx=16, ny=8, nz=16, substeps=120)
    sched, state = prepare(gpu, sc)
    runs = []
    for tune in (False, True):
        be = gpu.backend(F, sched, 0)
        try:
            be.upload(state, sc.settings, sc.sphere)
            s1 = be.run(control(50))
            if tune:
                before = be.status()
                be.autotune(8, 16)
                after = be.status()
                assert after.total_steps == before.total_steps == s1.total_steps
                assert after.broken_bonds == before.broken_bonds
            s2 = be.run(control(70))
            out = state.copy()
            sp = be.download(out)
            runs.append((s2, out, sp, be.fragment_labels()))
        finally:
            be.close()
    (a, oa, pa, la), (b, ob, pb, lb) = runs
    assert a.total_steps == b.total_steps == 120 and a.broken_bonds == b.broken_bonds > 0
    assert a.removed_energy_j == b.removed_energy_j and a.max_damage == b.max_damage
    for name in ("u", "v", "u_prev", "alive", "damage", "failure_mode", "prev_tensile"):
        assert np.array_equal(getattr(oa, name), getattr(ob, name)), name
    assert list(pa.center) == list(pb.center)
    assert la[1] == lb[1] and np.array_equal(la[0], lb[0])


def test_batch_autotune_changes_no_result(gpu):
    trials = [scenes.config3_trial(t, substeps=96) for t in range(5)]
    plain = run_batch(gpu, trials, F, steps_per_launch=24)
    tuned = run_batch(gpu, trials, F, steps_per_launch=24, autotune=(8, 16))
    for a, b in zip(tuned, plain):
        assert_same(a, b)
    assert tuned[0]["autotune"]["chosen"] in [
        {k: v for k, v in t.items() if k != "us_per_substep"} for t in tuned[0]["autotune"]["tried"]]


def test_tuning_is_validated(gpu):
    sc = scenes.small_impact(nx=8, ny=6, nz=8, substeps=4)
    sched, state = prepare(gpu, sc)
    be = gpu.backend(F, sched, 0)
    try:
        for bad in ({"colours_per_launch": 3}, {"lds_threads": 100}, {"graph_substeps": 1000}, {"streams": 3}):
            with pytest.raises(Exception):
                be.set_tuning(**bad)
        assert be.tuning() == {"colours_per_launch": 2, "lds_sweep": 1, "lds_threads": 1024, "graph_substeps": 0,
                               "streams": 0}
    finally:
        be.close()
