"""A fragment re-entering the lattice lane, on the device.

The island sub-lattice of the largest fragment of a shattered impact
(blx_fragment_lattice, equal to the reference's buildFragmentLattice:
tests/test_fragment_lattice.py), with its frozen deformation, turned pose and
spin velocities, is scheduled, built and run on the HIP backend; it equals the
oracle's run bit for bit, labels included.
"""
import pytest

from banjo_amd import abi
from banjo_amd.runner import run_scene
from test_fragment_lattice import island, island_scene
from test_gpu_parity import assert_same

pytestmark = pytest.mark.gpu


def test_island_runs_on_the_device_as_on_the_oracle(gpu, oracle):
    from banjo_amd import scenes
    sc = scenes.small_impact(nx=24, ny=12, nz=24, substeps=150)
    res = run_scene(oracle, sc, abi.PRECISION_FLOAT)
    sub, info = island(gpu, sc, res, 0)
    isc = island_scene(sc, sub, info, 120)
    print(f"island: {sub.node_count} cells, {sub.bond_count} bonds, 120 substeps", flush=True)
    h = run_scene(gpu, isc, abi.PRECISION_FLOAT, labels="device")
    o = run_scene(oracle, isc, abi.PRECISION_FLOAT)
    assert h["status"].total_steps == 120
    assert_same(h, o)
