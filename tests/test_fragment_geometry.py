"""The fragment build (rigid fragments and debris) against the reference's own.

banjo_amd/fragments.py composes the reference's buildFragmentRepresentations
(src/fracture/FragmentGeometry.cpp:203-291) from canonical labels, per-fragment
mass properties and the product library's geometry entry points
(blx_exposed_voxel_surface, blx_collision_points, blx_merge_cells_into_boxes).
On the CPU, with the oracle's mass properties (bit for bit the reference's
calculateFragmentMassProperties, tests/test_fragment_mass.py), the whole build
must equal the reference's function compiled into oracle/_ref bit for bit: the
order and split of rigid fragments and debris, every surface vertex, normal
and index, the collision proxies, voxel centres, debris radii and the momentum
totals. The box merge must equal mergeCellsIntoBoxes (PieceGeometry.cpp:9-46)
on the pieces' cells and on hand-made shapes.
"""
import ctypes as C

import numpy as np
import pytest

from banjo_amd import abi, fragments, scenes
from banjo_amd.lattice import _ptr
from banjo_amd.runner import run_scene

F = abi.PRECISION_FLOAT
DENSITY = 1000.0  # 1 g per 1 cm cell


def grid_of(sc):
    nx, ny, nz = sc.dims
    i = np.arange(sc.node_count)
    return np.stack([i % nx, (i // nx) % ny, i // (nx * ny)], axis=1).astype(np.int32)


def reference_build(ref, sc, pos, vel, labels, count, settings):
    N = sc.node_count
    g = np.ascontiguousarray(grid_of(sc)).reshape(-1)
    p = np.ascontiguousarray(pos, np.float64).reshape(-1)
    v = np.ascontiguousarray(vel, np.float64).reshape(-1)
    m = np.ascontiguousarray(sc.asset.mass, np.float64)
    lab = np.ascontiguousarray(labels, np.uint32)
    st = settings.as_array()
    size = C.c_uint64()
    fn = ref.fn("fragment_build")
    args = [C.c_uint32(N), _ptr(g, C.c_int32), _ptr(p, C.c_double), _ptr(v, C.c_double), _ptr(m, C.c_double), None,
            C.c_double(sc.cell), C.c_double(DENSITY), _ptr(lab, C.c_uint32), C.c_uint32(count), _ptr(st, C.c_double)]
    ref.check(fn(*args, None, C.c_uint64(0), C.byref(size)), "fragment_build")
    out = np.zeros(size.value)
    ref.check(fn(*args, _ptr(out, C.c_double), C.c_uint64(size.value), C.byref(size)), "fragment_build")
    return out


def reference_boxes(ref, cells):
    c = np.ascontiguousarray(cells, np.int32).reshape(-1)
    n = C.c_uint64()
    fn = ref.fn("merge_cells_into_boxes")
    ref.check(fn(_ptr(c, C.c_int32), C.c_uint64(len(c) // 3), None, C.c_uint64(0), C.byref(n)), "merge")
    out = np.zeros((n.value, 6), np.int32)
    ref.check(fn(_ptr(c, C.c_int32), C.c_uint64(len(c) // 3), _ptr(out, C.c_int32), C.c_uint64(n.value),
                 C.byref(n)), "merge")
    return out


@pytest.fixture(scope="module")
def shattered(oracle):
    sc = scenes.small_impact(nx=24, ny=12, nz=24, substeps=150)
    res = run_scene(oracle, sc, F)
    pos = sc.asset.reference_position + res["u"]
    return sc, res, pos


@pytest.mark.parametrize("settings", [fragments.FragmentBuildSettings(),
                                      fragments.FragmentBuildSettings(first_body_id=7, maximum_rigid_fragments=5,
                                                                      minimum_nodes_per_rigid_fragment=40,
                                                                      maximum_collision_points=8)],
                         ids=["defaults", "few_rigid"])
def test_fragment_build_equals_the_reference(hip, oracle, ref, shattered, settings):
    sc, res, pos = shattered
    labels, count = res["labels"], res["components"]
    assert count > 100  # rigid fragments and debris both
    order = np.argsort(labels, kind="stable")
    offsets = np.searchsorted(labels[order], np.arange(count + 1)).astype(np.uint32)
    mp = oracle.fragment_mass(pos, res["v"], sc.asset.mass, sc.cell, offsets, order.astype(np.uint32))
    mine = fragments.build_fragments(hip, labels, mp, grid_of(sc), pos, sc.cell, DENSITY, settings)
    theirs = reference_build(ref, sc, pos, res["v"], labels, count, settings)
    assert len(mine.rigid) == min(settings.maximum_rigid_fragments,
                                  max(1, int((np.bincount(labels) >= settings.minimum_nodes_per_rigid_fragment).sum())))
    assert len(mine.debris) > 0
    flat = mine.flatten()
    assert flat.shape == theirs.shape
    assert np.array_equal(flat, theirs)


def test_surface_of_a_block_and_of_a_single_cell(hip):
    # A 2x3x4 block has 2(6 + 8 + 12) = 52 exposed faces; one cell has six.
    g = np.array([(x, y, z) for z in range(4) for y in range(3) for x in range(2)], np.int32)
    p = g * 0.01
    vp, vn, idx = fragments.exposed_surface(hip, g, p, p.mean(0), 0.01)
    assert len(idx) == 6 * 52 and len(vp) == 4 * 52
    # Each face's normal points away from the block's centre.
    centre = p.mean(0) - p.mean(0)
    assert np.all(((vp - centre) * vn).sum(1) > 0)
    vp1, vn1, idx1 = fragments.exposed_surface(hip, g[:1], p[:1], p[0], 0.01)
    assert len(idx1) == 36
    assert np.allclose(np.abs(vp1), 0.005)


def test_box_merge_equals_the_reference(hip, ref, shattered):
    sc, res, pos = shattered
    g = grid_of(sc)
    labels = res["labels"]
    shapes = [g[labels == f] for f in range(6)]
    rng = np.random.default_rng(5)
    blob = np.array([(x, y, z) for z in range(5) for y in range(5) for x in range(5)], np.int32)
    shapes.append(blob[rng.random(len(blob)) < 0.6])
    shapes.append(np.array([(0, 0, 0), (-3, 2, 1), (-2, 2, 1), (5, -1, 0)], np.int32))
    for cells in shapes:
        mine = fragments.merge_cells_into_boxes(hip, cells)
        assert np.array_equal(mine, reference_boxes(ref, cells))
        # Every cell covered exactly once.
        vol = ((mine[:, 3:] - mine[:, :3] + 1).prod(1)).sum()
        assert vol == len(np.unique(cells, axis=0))
    solid = fragments.merge_cells_into_boxes(hip, blob)
    assert solid.tolist() == [[0, 0, 0, 4, 4, 4]]
