"""The C-ABI library loads without a GPU and exports every entry point the header declares."""
import ctypes
import os
import re

from conftest import HIP_LIB, ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "banjo_lattice.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(blx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_lattice_api():
    names = declared_functions()
    for required in ("blx_backend_create", "blx_backend_upload", "blx_backend_run", "blx_backend_download",
                     "blx_backend_set_external_forces", "blx_schedule_build", "blx_state_build",
                     "blx_backend_fragment_labels", "blx_components_host", "blx_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol(hip):
    lib = ctypes.CDLL(HIP_LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_surface_as_codes_and_messages(hip):
    lib = hip.lib
    lib.blx_backend_create.restype = ctypes.c_int
    out = ctypes.c_void_p()
    rc = lib.blx_backend_create(ctypes.c_int(99), ctypes.c_int(0), None, ctypes.byref(out))
    assert rc == -2  # null schedule: invalid argument
    assert b"null" in hip._err()
