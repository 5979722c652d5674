"""The C-ABI library loads and exports every entry point include/pdx.h declares (CPU only:
no call reaches the device)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(pdx_[a-z0-9_]+)\s*\(", txt))
    return names


def test_header_declares_entry_points():
    names = declared()
    for n in ("pdx_nl_create", "pdx_nl_update", "pdx_pair_forces", "pdx_mesh_spread", "pdx_mesh_solve",
              "pdx_mesh_gather", "pdx_ewald_fp64"):
        assert n in names


def test_library_exports_every_declared_symbol():
    import pdx
    lib = ctypes.CDLL(pdx.LIB_PATH)
    missing = [n for n in sorted(declared()) if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.pdx_version() >= 100


def test_argument_errors_without_device():
    """Host-side validation fails before any device call."""
    import pdx
    spec = pdx._lib.pdx_pair_spec()
    spec.model = 9
    h = ctypes.c_void_p()
    assert pdx.lib.pdx_pair_create(ctypes.byref(spec), ctypes.byref(h)) == pdx._lib.PDX_ERR_ARGUMENT
    assert "unknown pair model" in pdx.last_error()
    box = pdx._lib.dbl3((10.0, 10.0, 10.0))
    # fewer than 3 cells per dimension is rejected (NeighborAllocation.jl:15-20)
    assert pdx.lib.pdx_nl_create(10, box, 3.0, 0.5, 8, 20, 0, ctypes.byref(h)) == pdx._lib.PDX_ERR_ARGUMENT
    assert "3 cells" in pdx.last_error()
    assert pdx.lib.pdx_nl_create(0, box, 1.0, 0.5, 8, 20, 0, ctypes.byref(h)) == pdx._lib.PDX_ERR_ARGUMENT
    assert pdx.lib.pdx_nl_create(10, box, 1.0, 0.5, 8, 0, 0, ctypes.byref(h)) == pdx._lib.PDX_ERR_ARGUMENT
