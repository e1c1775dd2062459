"""The allocation-free fragment labels against the allocating entry point.

blx_backend_fragment_labels_noalloc runs the label kernels in one fixed
scratch array per GPU and writes into the caller's device buffer. Its labels
and component count must equal blx_backend_fragment_labels' bit for bit, on
two calls in a row (the second on a smaller lattice, so stale scratch from the
first call would show), and a lattice above its node limit is refused.
"""
import ctypes as C

import numpy as np
import pytest

from banjo_amd import abi, scenes
from banjo_amd.lattice import Asset
from banjo_amd.runner import prepare, run_scene

pytestmark = pytest.mark.gpu
MAX_NODES = 1 << 20  # the limit include/banjo_lattice.h states
HIP_DEVICE_TO_HOST = 2


def hip_runtime():
    """The HIP runtime the lattice library links (the process's only one)."""
    rt = C.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    rt.hipFree.argtypes = [C.c_void_p]
    return rt


def labels_noalloc(gpu, backend, node_count, rt):
    fn = gpu.fn("backend_fragment_labels_noalloc")
    buf = C.c_void_p()
    assert rt.hipMalloc(C.byref(buf), max(4, 4 * node_count)) == 0
    try:
        count = C.c_uint32()
        gpu.check(fn(backend.h, C.c_void_p(buf.value), C.byref(count)), "fragment_labels_noalloc")
        out = np.zeros(node_count, np.uint32)
        assert rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(buf.value), 4 * node_count,
                            HIP_DEVICE_TO_HOST) == 0
        return out, count.value
    finally:
        rt.hipFree(C.c_void_p(buf.value))


def test_noalloc_labels_equal_the_allocating_labels(gpu):
    rt = hip_runtime()
    cases = [scenes.small_impact(nx=24, ny=12, nz=24, substeps=150), scenes.small_impact(nx=12, ny=8, nz=12, substeps=90)]
    for sc in cases:  # the second call is on the smaller lattice
        print(f"no-alloc labels: {sc.name}, {sc.node_count} nodes", flush=True)
        res = run_scene(gpu, sc, abi.PRECISION_FLOAT, labels="device", keep_backend=True)
        be = res["backend"]
        try:
            got, count = labels_noalloc(gpu, be, sc.node_count, rt)
            assert count == res["components"] and count > 1
            assert np.array_equal(got, res["labels"])
            again, count2 = labels_noalloc(gpu, be, sc.node_count, rt)  # scratch holds the last call's contents
            assert count2 == count and np.array_equal(again, got)
        finally:
            be.close()


def test_noalloc_labels_refuse_a_lattice_above_the_limit(gpu):
    n = MAX_NODES + 1
    pos = np.zeros((n, 3))
    pos[:, 0] = 0.01 * np.arange(n)
    a = np.arange(0, 8, dtype=np.uint32)
    asset = Asset(reference_position=pos, mass=np.full(n, 1e-3), bond_a=a, bond_b=a + 1,
                  rest_length=np.full(8, 0.01), compliance=np.full(8, 1e-6), threshold=np.full((8, 6), np.inf))
    sc = scenes.Scene(name="chain", dims=(n, 1, 1), cell=0.01, asset=asset, slabs=np.array([0, n], np.uint32),
                      settings=scenes.settings(1e-4), sphere=scenes.sphere_state(center=(0.0, 100.0, 0.0)),
                      substeps=1)
    print(f"no-alloc labels: {n} nodes, above the limit", flush=True)
    sched, state = prepare(gpu, sc)
    be = gpu.backend(abi.PRECISION_FLOAT, sched, 0)
    try:
        be.upload(state, sc.settings, sc.sphere)
        count = C.c_uint32()
        rc = gpu.fn("backend_fragment_labels_noalloc")(be.h, C.c_void_p(16), C.byref(count))
        assert rc == -2 and b"at most" in gpu._err()
    finally:
        be.close()
