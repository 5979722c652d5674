"""ctypes binding of libpdx (include/pdx.h) and the error mapping of its status codes.

The library is loaded from ``particledynamics.jl_amd/lib/libpdx.so`` (built in-tree by
``__graft_entry__.build()`` / ``make -C particledynamics.jl_amd/csrc``). There is no CPU
fallback: importing this module without the built library raises ``ImportError``.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libpdx.so")

PDX_OK = 0
PDX_ERR_ARGUMENT = 1
PDX_ERR_DIMENSION = 2
PDX_ERR_CAPACITY = 3
PDX_ERR_COORDINATE = 4
PDX_ERR_DEVICE = 5
PDX_ERR_STATE = 6

PDX_DN_NSTATE = 23
PDX_NL_ALLPAIRS = 1
PDX_NL_DIAGNOSTICS = 2
PDX_NL_2D = 4
PDX_PAIR_NONE, PDX_PAIR_LJ, PDX_PAIR_WCA, PDX_PAIR_SOFT = 0, 1, 2, 3
PDX_PAIR_MIXED_LJ, PDX_PAIR_MIXED_WCA, PDX_PAIR_MATRIX_LJ, PDX_PAIR_MATRIX_WCA = 4, 5, 6, 7
PDX_ENERGY = 1
PDX_ADD = 2
PDX_INFLUENCE_SPME = 1
PDX_INFLUENCE_P3M_AD = 2
PDX_INFLUENCE_P3M_IK = 3
PDX_INFLUENCE_ESP = 4
PDX_INFLUENCE_ESP_DECONVOLUTION = 5


class ArgumentError(ValueError):
    """Invalid argument (the reference's ``ArgumentError``)."""


class DimensionMismatch(ValueError):
    """Array sizes disagree (the reference's ``DimensionMismatch``)."""


class DeviceError(RuntimeError):
    """A HIP runtime call failed."""


class NeighborCapacityError(RuntimeError):
    """A build required more row entries than ``cap`` (src/neighbors/NeighborTypes.jl:41-50)."""

    def __init__(self, capacity: int, required: int, msg: str = ""):
        self.capacity = int(capacity)
        self.required = int(required)
        super().__init__(msg or f"Neighbor capacity {capacity} is insufficient; at least {required} "
                                f"entries per particle are required. Rebuild with a larger cap.")


class pdx_nl_arrays(C.Structure):
    _fields_ = [("neighbors_flat", C.c_void_p), ("counts", C.c_void_p), ("cell_offsets", C.c_void_p),
                ("particle_ids_sorted", C.c_void_p), ("cell_of_particle", C.c_void_p),
                ("rref_x", C.c_void_p), ("rref_y", C.c_void_p), ("rref_z", C.c_void_p),
                ("packed", C.c_void_p), ("state", C.c_void_p)]


class pdx_pair_spec(C.Structure):
    _fields_ = [("model", C.c_int32), ("shift", C.c_int32), ("epsilon", C.c_double), ("sigma", C.c_double),
                ("cutoff", C.c_double), ("coulomb", C.c_int32), ("coulomb_shift", C.c_int32),
                ("coulomb_C", C.c_double), ("coulomb_cutoff", C.c_double), ("alpha", C.c_double),
                ("lanes", C.c_int32), ("split_shape", C.c_double)]


class pdx_mesh_spec(C.Structure):
    _fields_ = [("coulomb_prefactor", C.c_double), ("relative_permittivity", C.c_double),
                ("real_cutoff", C.c_double), ("alpha", C.c_double), ("real_space_tolerance", C.c_double),
                ("order", C.c_int32), ("mesh", C.c_int32 * 3), ("mesh_spacing", C.c_double),
                ("influence", C.c_int32), ("net_charge", C.c_int32), ("net_charge_tolerance", C.c_double),
                ("real_space_modifier", C.c_int32), ("alias_order", C.c_int32), ("split_shape", C.c_double),
                ("window_shape", C.c_double)]


def _load():
    # torch first: it bundles the HIP runtime (same soname as ROCm's). Loading libpdx first
    # would map a second runtime into the process; with torch loaded, libpdx binds to torch's,
    # so torch streams and allocations are valid handles for the library.
    import torch  # noqa: F401
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"libpdx is not built: {LIB_PATH} is missing (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    vp, i32, dbl = C.c_void_p, C.c_int32, C.c_double
    P = C.POINTER
    sig = {
        "pdx_version": ([], C.c_int),
        "pdx_last_error": ([], C.c_char_p),
        "pdx_nl_create": ([i32, P(dbl), dbl, dbl, i32, i32, i32, P(vp)], C.c_int),
        "pdx_nl_destroy": ([vp], C.c_int),
        "pdx_nl_info": ([vp, P(i32)], C.c_int),
        "pdx_nl_arrays_get": ([vp, P(pdx_nl_arrays)], C.c_int),
        "pdx_nl_update": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "pdx_nl_rebuild": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "pdx_nl_advance": ([vp, vp], C.c_int),
        "pdx_nl_rearm": ([vp, vp], C.c_int),
        "pdx_nl_resize_capacity": ([vp, i32], C.c_int),
        "pdx_nl_set_box": ([vp, P(dbl), vp], C.c_int),
        "pdx_nl_status": ([vp, P(C.c_int64), i32, vp], C.c_int),
        "pdx_pair_create": ([P(pdx_pair_spec), P(vp)], C.c_int),
        "pdx_pair_destroy": ([vp], C.c_int),
        "pdx_pair_info": ([vp, P(dbl)], C.c_int),
        "pdx_pair_set_lanes": ([vp, i32], C.c_int),
        "pdx_nl_set_row_lanes": ([vp, i32], C.c_int),
        "pdx_pair_forces": ([vp, vp, vp, vp, vp, vp, vp, vp, i32, vp], C.c_int),
        "pdx_pair_compute": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, vp], C.c_int),
        "pdx_pair_set_sigmas": ([vp, i32, P(C.c_float)], C.c_int),
        "pdx_pair_set_types": ([vp, i32, i32, P(i32), P(dbl), P(dbl), P(dbl)], C.c_int),
        "pdx_pair_cutoff": ([vp, P(dbl)], C.c_int),
        "pdx_mesh_create": ([P(pdx_mesh_spec), i32, P(dbl), P(vp)], C.c_int),
        "pdx_mesh_destroy": ([vp], C.c_int),
        "pdx_mesh_info": ([vp, P(dbl)], C.c_int),
        "pdx_esp_rule": ([dbl, i32, P(dbl)], C.c_int),
        "pdx_mesh_set_charges": ([vp, P(C.c_float), i32], C.c_int),
        "pdx_mesh_set_box": ([vp, P(dbl)], C.c_int),
        "pdx_mesh_spread": ([vp, vp, vp, vp, vp, vp, vp], C.c_int),
        "pdx_mesh_spread_cells": ([vp, vp, vp], C.c_int),
        "pdx_mesh_spread_snapshot": ([vp, vp, vp], C.c_int),
        "pdx_mesh_set_transform_threads": ([vp, i32], C.c_int),
        "pdx_mesh_set_spread_threads": ([vp, i32], C.c_int),
        "pdx_nl_check": ([vp, vp, vp, vp, vp, vp], C.c_int),
        "pdx_nl_build": ([vp, vp], C.c_int),
        "pdx_nl_spatial_order": ([vp, vp, vp], C.c_int),
        "pdx_nl_invalidate": ([vp, vp], C.c_int),
        "pdx_nl_check_kick_drift": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, dbl, dbl, vp, vp], C.c_int),
        "pdx_nve_kick": ([vp, vp, vp, vp, vp, vp, vp, dbl, dbl, vp, vp], C.c_int),
        "pdx_nl_snapshot": ([vp, vp], C.c_int),
        "pdx_mesh_solve": ([vp, i32, vp], C.c_int),
        "pdx_mesh_gather": ([vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, i32, i32, vp], C.c_int),
        "pdx_mesh_gather_cells": ([vp, vp, vp, vp, vp, vp, vp, i32, i32, vp], C.c_int),
        "pdx_mesh_totals": ([vp, P(dbl), vp], C.c_int),
        "pdx_mesh_copy_grid": ([vp, i32, vp, vp], C.c_int),
        "pdx_mesh_fft": ([vp, i32, vp, vp, vp], C.c_int),
        "pdx_ewald_fp64": ([i32, vp, vp, vp, vp, P(dbl), dbl, dbl, dbl, i32, i32, vp, vp, i32, P(dbl), vp],
                           C.c_int),
        "pdx_ewald_fp64_noalloc": ([i32, vp, vp, vp, vp, P(dbl), dbl, dbl, dbl, i32, i32, vp, vp, i32, P(dbl), vp],
                                   C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    return lib


lib = _load()


def last_error() -> str:
    m = lib.pdx_last_error()
    return m.decode() if m else ""


def check(rc: int, what: str = "") -> None:
    """Raise the Python mirror of a non-zero libpdx status."""
    if rc == PDX_OK:
        return
    msg = last_error() or what
    if rc == PDX_ERR_ARGUMENT or rc == PDX_ERR_COORDINATE or rc == PDX_ERR_STATE:
        raise ArgumentError(msg)
    if rc == PDX_ERR_DIMENSION:
        raise DimensionMismatch(msg)
    if rc == PDX_ERR_DEVICE:
        raise DeviceError(msg)
    raise RuntimeError(f"libpdx status {rc}: {msg}")


def dbl3(v):
    return (C.c_double * 3)(*[float(a) for a in v])


def ptr(t):
    """Device pointer of a torch tensor, a raw address or a c_void_p (None -> NULL)."""
    if t is None:
        return None
    if isinstance(t, C.c_void_p):
        return t
    if isinstance(t, int):
        return C.c_void_p(t)
    return C.c_void_p(t.data_ptr())


def stream_handle(stream=None):
    """hipStream_t of a torch stream (None -> torch's current stream)."""
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def adopt(owner, handle, destroy):
    """Destroy ``handle`` with ``destroy`` when ``owner`` is collected, or at interpreter exit
    while the library is still loaded (weakref.finalize runs its exit hook before modules are
    torn down, so no handle outlives the process unreleased)."""
    return weakref.finalize(owner, destroy, handle)
