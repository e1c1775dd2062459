"""The reference-side adapter, compiled in the reference's own terms.

banjo_amd/adapter/hip_lattice_backend.cpp implements the reference's
LatticeBackend and its four GPU-lane factory functions (FastLattice.hpp:401-406)
over include/banjo_lattice.h. oracle/Makefile compiles it against the
reference's headers and links it with the reference's own objects into two
check programs (oracle/adapter_check.cpp):
- with the product library and no device: the factory throws
  std::runtime_error and reports no device, as the stub it replaces does
  (CudaLatticeUnavailable.cpp:11-25), and every value type (StepSettings,
  SphereState, RunControl, RunStatus) crosses the C views field for field;
- with the same entry points over the oracle (oracle/adapter_shim.cpp): the
  adapter driven through LatticeBackend& -- upload, external loads, runs
  continued across calls, frames, download, first failures, status -- gives
  makeCpuLatticeBackend's answers bit for bit, in float and double, on one
  block and three slabs;
- with the product library on an MI355X (marked gpu): the HIP backend driven
  through the adapter gives makeCpuLatticeBackend's answers bit for bit.
The programs exist where the reference tree was present at build time.
"""
import os
import subprocess

import pytest

from conftest import ROOT

PRODUCT = os.path.join(ROOT, "oracle", "_ref", "adapter_check_product")
ORACLE = os.path.join(ROOT, "oracle", "_ref", "adapter_check_oracle")


def run(path, *args):
    if not os.path.exists(path):
        pytest.skip("adapter check not built (needs the reference tree at build time)")
    r = subprocess.run([path, *args], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("adapter_check ok")
    return r.stdout


def test_adapter_without_a_device(hip):
    if hip.lib.blx_available():
        pytest.skip("a device is visible: the device run is the gpu test")
    out = run(PRODUCT)
    assert "factory throws std::runtime_error" in out


def test_adapter_drives_the_oracle_as_the_reference_cpu_backend():
    out = run(ORACLE, "oracle")
    assert out.count("identical to makeCpuLatticeBackend") == 6


@pytest.mark.gpu
def test_adapter_drives_the_device_as_the_reference_cpu_backend(gpu):
    out = run(PRODUCT)
    assert out.count("identical to makeCpuLatticeBackend") == 6
