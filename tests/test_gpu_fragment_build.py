"""The fragment build driven from the device, against the reference's own.

A shattered impact runs on the HIP backend; its canonical labels and every
fragment's mass properties come from the device (blx_backend_fragment_labels,
blx_backend_fragment_mass_properties), the geometry from the product library,
composed by banjo_amd/fragments.py. The reference's buildFragmentRepresentations
(compiled into oracle/_ref) builds the same state. The structure must be
identical: how many fragments are rigid and how many debris, their order,
body ids, cell and face counts, every normal and index, the collision proxies'
sizes. The device sums each fragment's mass terms in a fixed tree order rather
than the reference's node order (tests/test_gpu_fragment_mass.py), so values
derived from the centre of mass agree to rounding: within 1e-9 relative and
1e-12 absolute (positions are centimetres).
"""
import numpy as np
import pytest

from banjo_amd import abi, fragments, scenes
from banjo_amd.runner import run_scene
from test_fragment_geometry import DENSITY, grid_of, reference_build

pytestmark = pytest.mark.gpu


def test_device_fragment_build_matches_the_reference(gpu, ref):
    sc = scenes.small_impact(nx=24, ny=12, nz=24, substeps=150)
    print(f"fragment build from the device: {sc.name}", flush=True)
    res = run_scene(gpu, sc, abi.PRECISION_FLOAT, labels="device", keep_backend=True)
    be = res["backend"]
    try:
        mp = be.fragment_mass_properties(sc.cell)
    finally:
        be.close()
    labels, count = res["labels"], res["components"]
    assert len(mp) == count > 100
    pos = sc.asset.reference_position + res["u"]
    settings = fragments.FragmentBuildSettings()
    mine = fragments.build_fragments(gpu, labels, mp, grid_of(sc), pos, sc.cell, DENSITY, settings)
    theirs = reference_build(ref, sc, pos, res["v"], labels, count, settings)
    flat = mine.flatten()
    sizes = np.bincount(labels)
    rigid = min(settings.maximum_rigid_fragments, max(1, int((sizes >= settings.minimum_nodes_per_rigid_fragment).sum())))
    assert len(mine.rigid) == rigid and len(mine.debris) == count - rigid > 0
    assert flat.shape == theirs.shape
    np.testing.assert_allclose(flat, theirs, rtol=1e-9, atol=1e-12)
    # The integer parts of the records are exact.
    for f in mine.rigid:
        assert np.array_equal(f.indices[:6], [0, 1, 2, 0, 2, 3])
