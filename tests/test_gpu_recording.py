"""The recording of a HIP run against the recording of the same run on the
oracle: every field of the measurement report that is not a timing or a name
(wall and kernel seconds, launches, per-phase seconds, rates, the backend
name) and every frame of the playback (cell positions, fragment ids, bond
aliveness and damage, ball centre) are identical."""
import json

import numpy as np
import pytest

from banjo_amd import abi, recording, scenes

pytestmark = pytest.mark.gpu

TIMING = {"backend", "phase_seconds", "wall_s", "kernel_s", "launches", "bond_updates_per_s", "wall_total_s",
          "realtime_ratio", "recording_wall_s"}


def strip(tree):
    if isinstance(tree, dict):
        return {k: strip(v) for k, v in tree.items() if k not in TIMING}
    return tree


@pytest.mark.parametrize("name,make", [
    ("impact_float", lambda: scenes.small_impact(nx=14, ny=8, nz=14, substeps=90)),
    ("cut_float", lambda: scenes.plank_cut(20, 6, 5, speed=4.0, substeps=90, ledge_cells=4)),
])
def test_hip_recording_matches_the_oracles(gpu, oracle, tmp_path, name, make):
    print(f"recording {name}", flush=True)
    out = {}
    for lib in (gpu, oracle):
        sc = make()
        path = tmp_path / f"{lib.name}.json"
        m, frames, final = recording.record_run(lib, sc, capture_stride=15, max_frames=10, playback_path=str(path))
        out[lib.name] = (m, frames, json.loads(path.read_text()))
    (mh, fh, dh), (mo, fo, do) = out[gpu.name], out[oracle.name]
    assert strip(recording.measurements_tree(mh)) == strip(recording.measurements_tree(mo))
    assert mh.broken_bonds > 0 and mh.components >= 1
    assert len(fh) == len(fo) > 1
    for a, b in zip(fh, fo):
        for k in ("positions", "ball_center", "alive", "damage", "component_ids"):
            assert np.array_equal(a[k], b[k]), k
    assert dh["frames"] == do["frames"] and dh["bodies"] == do["bodies"]
