"""Cluster-pair nonbonded engine (host side of libpdx's ``pdx_cluster_*``).

Mirrors ``ClusterPairEngine`` (src/clusterpairs/Engine.jl:5-95): the short-range forces of a
``PairForce`` (LJ / WCA / soft, real-space Coulomb, Coulomb exceptions) over 8 x 8 cluster
tiles with Newton's third law. ``update`` = ``update_list!`` (device-decided rebuild when a
particle moved more than skin/2 since the build, prune when one moved more than
prune_buffer/2 since the last prune), ``forces`` = ``compute_forces!``, ``compute`` = both
(``compute!``), ``statistics`` = ``statistics`` (builds and prunes), ``ensure_capacity`` =
``ensure_capacity!`` (List.jl:434-450). Outputs have the layout of ``PairForce.compute``.
"""
from __future__ import annotations

import ctypes as C
import math
from collections import namedtuple

from . import _lib as L
from .neighbors import device_view

ClusterStatus = namedtuple("ClusterStatus", "super_clusters outer working builds prunes error required capacity")


class pdx_cluster_arrays(C.Structure):
    _fields_ = [("perm", C.c_void_p), ("slot_of", C.c_void_p), ("packed", C.c_void_p), ("n_outer", C.c_void_p),
                ("outer", C.c_void_p), ("control", C.c_void_p), ("order", C.c_void_p), ("working", C.c_void_p),
                ("n_working", C.c_void_p)]


def _sig(lib):
    vp, i32, f64 = C.c_void_p, C.c_int32, C.c_double
    P = C.POINTER
    for name, args in {
        "pdx_cluster_create": [i32, P(f64), vp, f64, f64, i32, i32, P(vp)],
        "pdx_cluster_destroy": [vp],
        "pdx_cluster_info": [vp, P(C.c_int64)],
        "pdx_cluster_arrays_get": [vp, P(pdx_cluster_arrays)],
        "pdx_cluster_update": [vp, vp, vp, vp, vp, vp],
        "pdx_cluster_forces": [vp, vp, vp, vp, vp, vp, vp, i32, vp],
        "pdx_cluster_request_build": [vp, vp],
        "pdx_cluster_status": [vp, P(C.c_int64), i32, vp],
        "pdx_cluster_resize": [vp, i32],
        "pdx_cluster_set_launch": [vp, i32, i32],
    }.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int


_sig(L.lib)


class ClusterPairEngine:
    """Cluster-pair list for ``n`` particles in a centred box, evaluating ``pair`` (a
    ``PairForce``, which must outlive the engine). ``prune_buffer`` defaults to half the skin
    (Engine.jl:50); ``capacity`` is the outer-entry capacity per super-cluster (0: estimated);
    ``waves`` the wave64s per super-cluster of the force pass."""

    def __init__(self, n, box, pair, skin, prune_buffer=None, capacity=0, waves=0):
        if prune_buffer is None:
            prune_buffer = 0.5 * float(skin)
        h = C.c_void_p()
        L.check(L.lib.pdx_cluster_create(int(n), L.dbl3(box), pair._h, float(skin), float(prune_buffer),
                                         int(capacity), int(waves), C.byref(h)))
        self._h = h
        self._release = L.adopt(self, h, L.lib.pdx_cluster_destroy)
        self.pair = pair  # keeps the model alive
        self.N = int(n)
        self.box = tuple(float(b) for b in box)
        self.skin = float(skin)
        self.prune_buffer = float(prune_buffer)
        self._refresh()

    def _refresh(self):
        o = (C.c_int64 * 10)()
        L.check(L.lib.pdx_cluster_info(self._h, o))
        keys = ("N", "ncx", "ncy", "nzb", "NSmax", "nslot", "capacity", "waves", "splits", "chunks")
        self.info = dict(zip(keys, [int(v) for v in o]))
        a = pdx_cluster_arrays()
        L.check(L.lib.pdx_cluster_arrays_get(self._h, C.byref(a)))
        self._arrays = a


    # ---- enqueue (never synchronises)
    def update(self, x, y, z, q=None, stream=None):
        L.check(L.lib.pdx_cluster_update(self._h, L.ptr(x), L.ptr(y), L.ptr(z), L.ptr(q), L.stream_handle(stream)))

    def forces(self, fx, fy, fz, Epot=None, V=None, totals=None, energy=False, add=False, stream=None):
        flags = (L.PDX_ENERGY if energy else 0) | (L.PDX_ADD if add else 0)
        L.check(L.lib.pdx_cluster_forces(self._h, L.ptr(fx), L.ptr(fy), L.ptr(fz), L.ptr(Epot), L.ptr(V),
                                         L.ptr(totals), flags, L.stream_handle(stream)))

    def compute(self, x, y, z, q=None, energy=True, stream=None):
        """update + forces into fresh arrays: (fx, fy, fz, Epot, V, totals), as PairForce.compute."""
        import torch
        n = self.N
        fx, fy, fz = (torch.empty(n, device="cuda") for _ in range(3))
        Epot = torch.empty(n, device="cuda") if energy else None
        V = torch.empty(6, n, device="cuda") if energy else None
        tot = torch.zeros(7, dtype=torch.float64, device="cuda") if energy else None
        self.update(x, y, z, q, stream=stream)
        self.forces(fx, fy, fz, Epot, V, tot, energy=energy, stream=stream)
        return fx, fy, fz, Epot, V, tot

    def request_build(self, stream=None):
        L.check(L.lib.pdx_cluster_request_build(self._h, L.stream_handle(stream)))

    # ---- host boundaries
    def status(self, check=True, stream=None):
        o = (C.c_int64 * 8)()
        rc = L.lib.pdx_cluster_status(self._h, o, 1 if check else 0, L.stream_handle(stream))
        if rc == L.PDX_ERR_CAPACITY:
            raise L.NeighborCapacityError(int(o[7]), int(o[6]), L.last_error())
        L.check(rc)
        return ClusterStatus(*[int(v) for v in o])

    def statistics(self):
        s = self.status(check=False)
        return {"builds": s.builds, "prunes": s.prunes, "outer": s.outer, "working": s.working,
                "super_clusters": s.super_clusters}

    def set_launch(self, waves, splits):
        """Force-pass launch shape: waves per workgroup and workgroups per super-cluster."""
        L.check(L.lib.pdx_cluster_set_launch(self._h, int(waves), int(splits)))
        self._refresh()

    def resize(self, capacity):
        L.check(L.lib.pdx_cluster_resize(self._h, int(capacity)))
        self._refresh()

    def ensure_capacity(self, growth=1.25):
        """After an outer-list overflow: grow the segments past the required size and request a
        build; True when it grew (the caller then updates and evaluates again)."""
        s = self.status(check=False)
        if s.error != 1:
            return False
        self.resize(int(math.ceil(growth * s.required)) + 16)
        return True

    # ---- device views (read-only)
    def perm(self):
        return device_view(self._arrays.perm, (self.info["nslot"],), _i32())

    def slot_of(self):
        return device_view(self._arrays.slot_of, (self.N,), _i32())

    def packed(self):
        import torch
        return device_view(self._arrays.packed, (self.info["nslot"], 4), torch.float32)

    def order(self):
        """1-based particle ids in slot order (the mesh traversal order)."""
        return device_view(self._arrays.order, (self.N,), _i32())

    @property
    def order_ptr(self):
        return self._arrays.order

    def outer(self):
        """(n_outer per super-cluster, entries (NSmax, capacity)): entry = (cj << 8) | i-mask."""
        return (device_view(self._arrays.n_outer, (self.info["NSmax"],), _i32()),
                device_view(self._arrays.outer, (self.info["NSmax"], self.info["capacity"]), _i32()))


    def tile_counts(self):
        """Host counts of the current lists: outer and working entries, and the 8 x 8 tiles
        (mask bits) of each."""
        import numpy as np
        NS = self.status(check=False).super_clusters
        ne, ent = (t.cpu().numpy() for t in self.outer())
        K, cap = self.info["chunks"], self.info["capacity"]
        went = device_view(self._arrays.working, (self.info["NSmax"], cap), _i32()).cpu().numpy()
        nw = device_view(self._arrays.n_working, (self.info["NSmax"] * K,), _i32()).cpu().numpy()
        bits = np.array([bin(k).count("1") for k in range(256)])
        out = dict(outer=0, outer_tiles=0, working=0, working_tiles=0)
        for S in range(NS):
            n = int(ne[S])
            out["outer"] += n
            out["outer_tiles"] += int(bits[ent[S, :n] & 255].sum())
            for w in range(K):
                e0 = (n * w) // K
                k = int(nw[S * K + w])
                out["working"] += k
                out["working_tiles"] += int(bits[went[S, e0:e0 + k] & 255].sum())
        return out


def _i32():
    import torch
    return torch.int32
