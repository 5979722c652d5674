"""Device cell list + slot-major ELL rows (host side of libpdx's neighbor list).

Mirrors ``DeviceNeighborWorkspace`` (src/neighbors/DeviceNeighborCUDA.jl:35-82):
``update`` = ``update_device_neighbors!`` (coverage check + predicated rebuild, no host
read), ``status`` = ``device_neighbor_status`` (the boundary readback that raises
``NeighborCapacityError`` / ``ArgumentError``), ``advance``, ``rearm`` and
``resize_capacity`` as in the reference. ``build_neighbors_dense`` mirrors
``NL.build_neighbors_dense!`` (src/neighbors/NeighborAllocation.jl:198-214).
"""
from __future__ import annotations

import ctypes as C
from collections import namedtuple

from . import _lib as L

NeighborStatus = namedtuple("NeighborStatus", "valid error requested step last_build interval checks builds required")

# 0-based state-word indices (DN_* of DeviceNeighborCUDA.jl:4-33, minus one)
DN = dict(REQUEST=0, VALID=1, ERROR=2, STEP=3, LAST_BUILD=4, INTERVAL=5, CHECKS=6, BUILDS=7, REQUIRED=8,
          REASON=9, ATTEMPTS=10, DISPLACEMENT_BUILDS=11, AGE_BUILDS=12, INVALID_BUILDS=13, SORT_BUILDS=14,
          OVERLAP_BUILDS=15, LAST_CANDIDATES=16, TOTAL_CANDIDATES=17, LAST_ACCEPTED=18, TOTAL_ACCEPTED=19,
          MAX_ROW=20, DIAGNOSTICS=21, SECTION=22)


class _CAI:
    """Minimal __cuda_array_interface__ holder to view library-owned device memory."""

    def __init__(self, ptr, shape, typestr):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (int(ptr), False),
                                         "version": 2, "strides": None}


def device_view(ptr, shape, dtype):
    import torch
    typestr = {torch.int32: "<i4", torch.float32: "<f4", torch.int64: "<i8", torch.float64: "<f8"}[dtype]
    return torch.as_tensor(_CAI(ptr, shape, typestr), device="cuda")


class DeviceNeighborWorkspace:
    """GPU-owned cell list and ELL rows for ``N`` particles in a centred orthorhombic box."""

    def __init__(self, N, box, cutoff, skin, cap, interval=20, allpairs=False, diagnostics=True):
        """``box`` of two lengths makes a 2-D list (DenseNeighborCUDA.jl:77): z may then be None."""
        h = C.c_void_p()
        if len(box) not in (2, 3):
            raise L.ArgumentError("box must have 2 or 3 lengths")
        self.dim = len(box)
        flags = ((L.PDX_NL_ALLPAIRS if allpairs else 0) | (L.PDX_NL_DIAGNOSTICS if diagnostics else 0)
                 | (L.PDX_NL_2D if self.dim == 2 else 0))
        box3 = tuple(box) + ((1.0,) if self.dim == 2 else ())
        L.check(L.lib.pdx_nl_create(int(N), L.dbl3(box3), float(cutoff), float(skin), int(cap), int(interval),
                                    flags, C.byref(h)))
        self._h = h
        self._release = L.adopt(self, h, L.lib.pdx_nl_destroy)
        self.N = int(N)
        self.box = tuple(float(b) for b in box)
        self.cutoff = float(cutoff)
        self.skin = float(skin)
        self._refresh_info()

    def _refresh_info(self):
        info = (C.c_int32 * 8)()
        L.check(L.lib.pdx_nl_info(self._h, info))
        self.cap, self.nx, self.ny, self.nz, self.ncell = info[1], info[2], info[3], info[4], info[5]
        self.allpairs = bool(info[6])
        a = L.pdx_nl_arrays()
        L.check(L.lib.pdx_nl_arrays_get(self._h, C.byref(a)))
        self._arrays = a

    @property
    def handle(self):
        return self._h


    # ---- enqueue (never synchronises)
    def update(self, x, y, z, q=None, stream=None):
        L.check(L.lib.pdx_nl_update(self._h, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(q), L.stream_handle(stream)))

    def rebuild(self, x, y, z, q=None, stream=None):
        L.check(L.lib.pdx_nl_rebuild(self._h, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(q), L.stream_handle(stream)))

    def check(self, x, y, z, q=None, stream=None):
        """First half of ``update``: coverage check and packing (no rebuild kernels)."""
        L.check(L.lib.pdx_nl_check(self._h, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(q), L.stream_handle(stream)))

    def check_kick_drift(self, x, y, z, q, vx, vy, vz, fx, fy, fz, dt, mass=1.0, inv_mass=None, stream=None):
        """``check`` preceded by the velocity-Verlet kick-drift (v += dt/(2m) f, x += dt v, wrapped),
        fused into the same pass; x, y, z, vx, vy, vz change in place."""
        L.check(L.lib.pdx_nl_check_kick_drift(self._h, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(q), L.ptr(vx), L.ptr(vy),
                                              L.ptr(vz), L.ptr(fx), L.ptr(fy), L.ptr(fz), float(dt), float(mass),
                                              L.ptr(inv_mass), L.stream_handle(stream)))

    def kick(self, vx, vy, vz, fx, fy, fz, dt, mass=1.0, inv_mass=None, stream=None):
        """Closing kick v += dt/(2m) f of a velocity-Verlet step; advances the step counter."""
        L.check(L.lib.pdx_nve_kick(self._h, L.ptr(vx), L.ptr(vy), L.ptr(vz), L.ptr(fx), L.ptr(fy), L.ptr(fz),
                                   float(dt), float(mass), L.ptr(inv_mass), L.stream_handle(stream)))

    def spatial_order(self, stream=None):
        """Permutation into the spatial order of the last build (``pdx_nl_spatial_order``): int64
        device tensor, slot -> 0-based particle id."""
        import torch
        perm = torch.empty(self.N, dtype=torch.int32, device="cuda")
        L.check(L.lib.pdx_nl_spatial_order(self._h, L.ptr(perm), L.stream_handle(stream)))
        return perm.long()

    def set_row_lanes(self, lanes: int):
        """Lanes per particle of the row build (the builder lanes of the engine settings)."""
        L.check(L.lib.pdx_nl_set_row_lanes(self._h, int(lanes)))
        self.row_lanes_set = int(lanes)

    def invalidate(self, stream=None):
        """Particles were relabelled: the next check rebuilds (``pdx_nl_invalidate``)."""
        L.check(L.lib.pdx_nl_invalidate(self._h, L.stream_handle(stream)))

    def build(self, stream=None):
        """Second half of ``update``: the rebuild kernels, predicated on the check's decision."""
        L.check(L.lib.pdx_nl_build(self._h, L.stream_handle(stream)))

    def snapshot(self, stream=None):
        """Copy the cells of a build completed since the last snapshot (for spread_snapshot); the
        check does the same, so this is only needed between checks."""
        L.check(L.lib.pdx_nl_snapshot(self._h, L.stream_handle(stream)))

    def advance(self, stream=None):
        L.check(L.lib.pdx_nl_advance(self._h, L.stream_handle(stream)))

    def rearm(self, stream=None):
        L.check(L.lib.pdx_nl_rearm(self._h, L.stream_handle(stream)))

    # ---- host boundaries
    def resize_capacity(self, cap):
        L.check(L.lib.pdx_nl_resize_capacity(self._h, int(cap)))
        self._refresh_info()

    def set_box(self, box, stream=None):
        if len(box) != self.dim:
            raise L.DimensionMismatch("the box has another dimension than the list")
        box3 = tuple(box) + ((1.0,) if self.dim == 2 else ())
        L.check(L.lib.pdx_nl_set_box(self._h, L.dbl3(box3), L.stream_handle(stream)))
        self.box = tuple(float(b) for b in box)
        self._refresh_info()

    def state(self, stream=None):
        s = (C.c_int64 * L.PDX_DN_NSTATE)()
        L.check(L.lib.pdx_nl_status(self._h, s, 0, L.stream_handle(stream)))
        return list(s)

    def status(self, check=True, stream=None):
        s = (C.c_int64 * L.PDX_DN_NSTATE)()
        rc = L.lib.pdx_nl_status(self._h, s, 1 if check else 0, L.stream_handle(stream))
        if rc == L.PDX_ERR_CAPACITY:
            raise L.NeighborCapacityError(self.cap, s[DN["REQUIRED"]], L.last_error())
        L.check(rc)
        return NeighborStatus(bool(s[1]), s[2], bool(s[0]), s[3], s[4], s[5], s[6], s[7], s[8])

    # ---- device views of the list storage (reference layouts)
    def neighbors_flat(self):
        """Slot-major rows, 1-based ids: element [t, i] is slot t of particle i (shape cap x N)."""
        import torch
        return device_view(self._arrays.neighbors_flat, (self.cap, self.N), torch.int32)

    def counts(self):
        import torch
        return device_view(self._arrays.counts, (self.N,), torch.int32)

    def cell_offsets(self):
        import torch
        return device_view(self._arrays.cell_offsets, (self.ncell + 1,), torch.int32)

    def particle_ids_sorted(self):
        import torch
        return device_view(self._arrays.particle_ids_sorted, (self.N,), torch.int32)

    def cell_of_particle(self):
        import torch
        return device_view(self._arrays.cell_of_particle, (self.N,), torch.int32)

    def packed(self):
        import torch
        return device_view(self._arrays.packed, (self.N, 4), torch.float32)

    def rows_host(self):
        """Rows as Python lists of 0-based ids (boundary helper for tests)."""
        counts = self.counts().cpu().numpy()
        flat = self.neighbors_flat().cpu().numpy()
        out = []
        for i in range(self.N):
            n = min(int(counts[i]), self.cap)
            out.append(sorted(int(v) - 1 for v in flat[:n, i]))
        return out


def build_neighbors_dense(x, y, z, box, cutoff, skin, cap, q=None, allpairs=False, stream=None):
    """Build rows now and check capacity (raises ``NeighborCapacityError``)."""
    ws = DeviceNeighborWorkspace(x.numel(), box, cutoff, skin, cap, allpairs=allpairs)  # len(box) 2: 2-D
    ws.rebuild(x, y, z, q, stream=stream)
    ws.status(check=True, stream=stream)
    return ws


CellList = DeviceNeighborWorkspace
