"""Writes the golden fixtures under tests/golden/ from the reference build.

The reference's lattice lane is compiled from its own sources into
oracle/_ref/libbanjo_ref.so (oracle/Makefile); this script drives it through
the same scene builders the tests use and stores inputs' seeds and the
outputs only (no reference code or text):

  c1_float.npz          config 1 cantilever, 16^3, full 1,920 substeps, float
  c1_double.npz         the same in double with absolute-position arithmetic
                        (the FP64 run positions are reported against)
  impact_4slabs.npz     a 24x12x24 impact block on 4 slabs, 150 substeps, float

Run:  python tests/golden/make_golden.py   (needs oracle/_ref built)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from banjo_amd import abi, scenes  # noqa: E402
from banjo_amd.lattice import LatticeLibrary  # noqa: E402
from banjo_amd.runner import run_scene  # noqa: E402

STATUS = ("total_steps", "first_failure_step", "last_failure_step", "failure_rounds", "broken_bonds",
          "removed_energy_j", "last_energy_gain_step", "max_damage", "last_damage_gain_step", "exit_reason")

CASES = {
    "c1_float": (lambda: scenes.config1(), abi.PRECISION_FLOAT, False),
    "c1_double": (lambda: scenes.config1(), abi.PRECISION_DOUBLE, True),
    "impact_4slabs": (lambda: scenes.small_impact(substeps=150, blocks=4), abi.PRECISION_FLOAT, False),
}


def main():
    ref = LatticeLibrary(os.path.join(ROOT, "oracle", "_ref", "libbanjo_ref.so"), "ref_", "ref")
    for name, (make, precision, direct) in CASES.items():
        sc = make()
        sc.settings.direct_arithmetic = 1 if direct else 0
        # The serial CpuLatticeBackend: the reference's GPU lane was proven equal to it
        # (src/fastlattice/FastLattice.hpp:7-11). The threaded backend matches it on
        # every array but books last_energy_gain_step per round rather than per bond
        # (ParallelCpuLatticeBackend.cpp:1914-1919 against CpuLatticeBackend.cpp:472-474).
        r = run_scene(ref, sc, precision, backend_args=(0, 1))
        st = r["status"]
        fdt = np.float32 if precision == abi.PRECISION_FLOAT else np.float64
        np.savez_compressed(
            os.path.join(HERE, name + ".npz"),
            u=r["u"].astype(fdt), v=r["v"].astype(fdt), alive=np.packbits(r["alive"]),
            mode=r["mode"], labels=r["labels"].astype(np.uint32), components=np.uint32(r["components"]),
            substeps=np.uint64(sc.substeps), precision=np.int32(precision), direct=np.int32(direct),
            sphere_center=np.array(r["sphere"].center[:]),
            **{f: np.array(getattr(st, f)) for f in STATUS})
        print(name, "broken", st.broken_bonds, "components", r["components"], "exit", st.exit_reason)


if __name__ == "__main__":
    main()
