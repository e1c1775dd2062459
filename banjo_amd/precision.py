"""Float against double on the HIP backend: the precision report.

The float lane is the product path. The double lane computes the same
expressions in FP64, and the HIP double backend equals the reference's own
double lane bit for bit (the oracle is pinned to the reference build, and the
HIP backend to the oracle, in both precisions), so the double run is the
reference's FP64 answer for the scene. For one scene both lanes run on the
device from the same state, and the report gives:

- max |du| (m) and max |dv| (m/s) of float against double, beside the largest
  |u| and |v| of the double run for scale. The reference's own bar for its
  float lane is max |du| < 1e-6 m over a small impact
  (tests/fast_lattice_tests.cpp:435-447);
- the bonds alive in one run and broken in the other, and how many of them
  lie within `eps` (relative) of their break threshold. A bond's distance from
  threshold is measured in the run in which it broke: the sample that broke it
  (a dead bond keeps the last sample, LatticePhysics.hpp:1069-1114) against
  the threshold of its failure mode's ramp end, where damage reaches 1. A bond
  that barely crossed in one lane and stayed just below in the other is a
  near-threshold disagreement; one far from threshold is a divergence of the
  trajectories (a fracture cascade amplifies rounding differences);
- whether the fragment labels and counts agree.
"""
from __future__ import annotations

import numpy as np

from . import abi
from .runner import run_scene

EPS = 1.0e-3  # relative distance from threshold that counts as near-threshold

# Failure mode -> (state field of its last sample, threshold column of its ramp end)
_MODE_SAMPLE = {1: ("prev_tensile", 1), 2: ("prev_compressive", 3), 3: ("prev_shear", 5)}


def _asset_order(res, name):
    arr = getattr(res["state"], name)
    out = np.empty_like(arr)
    out[res["schedule"].bond_order] = arr
    return out


def compare(scene, f, d, eps=EPS):
    """The report of a float result f against a double result d of `scene`
    (run_scene results with labels)."""
    du = np.abs(f["u"] - d["u"])
    dv = np.abs(f["v"] - d["v"])
    diff = np.nonzero(f["alive"] != d["alive"])[0]
    rel = np.zeros(len(diff))
    thr = scene.asset.threshold
    samples = {}
    for k, j in enumerate(diff):
        res = f if f["alive"][j] == 0 else d  # the run in which the bond broke
        mode = int(res["mode"][j])
        field, col = _MODE_SAMPLE.get(mode, ("prev_tensile", 1))
        key = (id(res), field)
        if key not in samples:
            samples[key] = _asset_order(res, field)
        t = thr[j, col]
        rel[k] = abs(samples[key][j] - t) / abs(t) if np.isfinite(t) and t != 0 else np.inf
    return {
        "workload": scene.name, "nodes": scene.node_count, "bonds": scene.bond_count,
        "substeps": int(d["status"].total_steps),
        "max_abs_du_m": float(du.max()), "max_abs_dv_m_s": float(dv.max()),
        "max_abs_u_m": float(np.abs(d["u"]).max()), "max_abs_v_m_s": float(np.abs(d["v"]).max()),
        "broken_float": int(f["status"].broken_bonds), "broken_double": int(d["status"].broken_bonds),
        "alive_mismatches": int(len(diff)), "eps": eps,
        "mismatches_within_eps": int((rel <= eps).sum()),
        "mismatches_beyond_eps": int((rel > eps).sum()),
        "max_relative_distance_from_threshold": float(rel.max()) if len(diff) else 0.0,
        "labels_equal": bool(np.array_equal(f["labels"], d["labels"])),
        "fragments_float": int(f["components"]), "fragments_double": int(d["components"]),
    }


def report(lib, scene, eps=EPS):
    """Runs `scene` on the HIP backend in float and in double and compares."""
    f = run_scene(lib, scene, abi.PRECISION_FLOAT, labels="device")
    d = run_scene(lib, scene, abi.PRECISION_DOUBLE, labels="device")
    out = compare(scene, f, d, eps)
    out["seconds_float"], out["seconds_double"] = f["wall"], d["wall"]
    return out
