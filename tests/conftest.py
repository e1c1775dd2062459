"""Shared fixtures: the product library and the checkers (oracle, reference build)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

ORACLE_LIB = os.path.join(ROOT, "oracle", "build", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libbanjo_ref.so")
HIP_LIB = os.path.join(ROOT, "banjo_amd", "lib", "libbanjo_lattice.so")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")


def _ensure_built():
    if not os.path.exists(HIP_LIB) or not os.path.exists(ORACLE_LIB):
        import __graft_entry__
        __graft_entry__.build()


@pytest.fixture(scope="session")
def oracle():
    _ensure_built()
    from banjo_amd.lattice import LatticeLibrary
    return LatticeLibrary(ORACLE_LIB, "orc_", "oracle")


@pytest.fixture(scope="session")
def ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("reference build oracle/_ref absent (needs the reference tree at build time)")
    from banjo_amd.lattice import LatticeLibrary
    return LatticeLibrary(REF_LIB, "ref_", "ref")


@pytest.fixture(scope="session")
def hip():
    _ensure_built()
    from banjo_amd.lattice import LatticeLibrary
    return LatticeLibrary(HIP_LIB, "blx_", "hip")


@pytest.fixture(scope="session")
def gpu(hip):
    """The HIP library on a machine with a gfx950 device; fails loudly otherwise."""
    if not hip.lib.blx_available():
        pytest.fail("no gfx950 device visible to the HIP lattice library")
    return hip
