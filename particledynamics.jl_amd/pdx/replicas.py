"""Batched replica boxes advanced by velocity Verlet in one persistent launch (host side of
libpdx's ``pdx_replicas_*``), and the launch-shape autotuner.

Each replica is an independent orthorhombic box of LJ or WCA atoms (unit mass by default)
that one ``Run`` with ``UniformLJ``/``UniformWCA`` and the NVE kick-drift / kick kernels
would advance (src/engine/Run.jl:247-294, src/simulation/NVEFusion.jl:3-24). All replicas'
atoms live in concatenated SoA device arrays; replica k occupies
``[offsets[k], offsets[k+1])``.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import time

import numpy as np

from . import _lib as L

PDX_VV_FORCES_VALID = 4


class pdx_replica_spec(C.Structure):
    _fields_ = [("model", C.c_int32), ("shift", C.c_int32), ("n", C.c_int32), ("reserved", C.c_int32),
                ("epsilon", C.c_double), ("sigma", C.c_double), ("cutoff", C.c_double), ("box", C.c_double * 3)]


def _sig(lib):
    vp, i32, f64 = C.c_void_p, C.c_int32, C.c_double
    P = C.POINTER
    for name, args in {
        "pdx_replicas_create": [i32, P(pdx_replica_spec), i32, i32, P(vp)],
        "pdx_replicas_destroy": [vp],
        "pdx_replicas_set_launch": [vp, i32, i32],
        "pdx_replicas_info": [vp, P(C.c_int64)],
        "pdx_replicas_offsets": [vp, P(C.c_int64)],
        "pdx_replicas_run": [vp] + [vp] * 9 + [f64, f64, i32, i32, vp, vp],
    }.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int


_sig(L.lib)

_MODELS = {"lj": L.PDX_PAIR_LJ, "wca": L.PDX_PAIR_WCA}


class ReplicaBatch:
    """``specs``: sequence of dicts with keys ``model`` ("lj" | "wca"), ``n``, ``box`` (3),
    and optionally ``epsilon`` (1), ``sigma`` (1), ``cutoff`` (2.5, LJ only), ``shift``
    (True: LJ energy zero at the cutoff)."""

    def __init__(self, specs, block=0, stencil=0):
        arr = (pdx_replica_spec * len(specs))()
        for k, s in enumerate(specs):
            m = s["model"]
            if m not in _MODELS:
                raise L.ArgumentError(f"replica {k}: model must be 'lj' or 'wca'")
            a = arr[k]
            a.model = _MODELS[m]
            a.shift = 1 if s.get("shift", True) else 0
            a.n = int(s["n"])
            a.epsilon = float(s.get("epsilon", 1.0))
            a.sigma = float(s.get("sigma", 1.0))
            a.cutoff = float(s.get("cutoff", 2.5))
            for d in range(3):
                a.box[d] = float(s["box"][d])
        h = C.c_void_p()
        L.check(L.lib.pdx_replicas_create(len(specs), arr, int(block), int(stencil), C.byref(h)))
        self._h = h
        self._release = L.adopt(self, h, L.lib.pdx_replicas_destroy)
        self.specs = list(specs)
        self.nrep = len(specs)
        off = (C.c_int64 * (self.nrep + 1))()
        L.check(L.lib.pdx_replicas_offsets(h, off))
        self.offsets = np.array(off[:], dtype=np.int64)
        self.total = int(self.offsets[-1])


    def info(self):
        o = (C.c_int64 * 8)()
        L.check(L.lib.pdx_replicas_info(self._h, o))
        keys = ("atoms", "replicas", "max_n", "lds_bytes", "block", "stencil", "allpairs_replicas", "max_cells")
        return dict(zip(keys, [int(v) for v in o]))

    def set_launch(self, block, stencil):
        L.check(L.lib.pdx_replicas_set_launch(self._h, int(block), int(stencil)))

    def run(self, state, nsteps, dt=0.005, mass=1.0, energy=False, forces_valid=True, out=None, stream=None):
        """Advance ``state`` (dict of device FP32 arrays x, y, z, vx, vy, vz, fx, fy, fz, each of
        ``total`` atoms) by ``nsteps`` steps in place. Returns the (nrep, 4) FP64 device array
        [potential, kinetic, sum r.F, n] when ``energy``."""
        import torch
        for k in ("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"):
            t = state[k]
            if t.numel() != self.total or t.dtype != torch.float32 or not t.is_cuda:
                raise L.DimensionMismatch(f"{k} must be a device float32 array of {self.total} atoms")
        if energy and out is None:
            out = torch.zeros(self.nrep, 4, dtype=torch.float64, device="cuda")
        flags = (L.PDX_ENERGY if energy else 0) | (PDX_VV_FORCES_VALID if forces_valid else 0)
        L.check(L.lib.pdx_replicas_run(self._h, *(L.ptr(state[k]) for k in ("x", "y", "z", "vx", "vy", "vz", "fx",
                                                                           "fy", "fz")),
                                       float(mass), float(dt), int(nsteps), flags, L.ptr(out),
                                       L.stream_handle(stream)))
        return out

    def forces(self, state, energy=True, stream=None):
        """Forces (and energies) of the current coordinates; no integration."""
        return self.run(state, 0, energy=energy, forces_valid=False, stream=stream)


def device_state(batch_dict):
    """Device copies (torch, FP32) of a host batch from ``replica_batch``."""
    import torch
    st = {k: torch.from_numpy(np.ascontiguousarray(batch_dict[k], dtype=np.float32)).cuda()
          for k in ("x", "y", "z", "vx", "vy", "vz")}
    for k in ("fx", "fy", "fz"):
        st[k] = torch.zeros_like(st["x"])
    return st


def clone_state(state):
    return {k: v.clone() for k, v in state.items()}


def autotune(batch: ReplicaBatch, state, steps=20, candidates=None, table=None, key=None, dt=0.005, mass=1.0):
    """Pick the fastest launch shape (workgroup size, cells per cutoff) for this batch by
    timing ``steps`` steps of each candidate on a scratch copy of ``state``; the winner is
    set on ``batch``. With ``table`` (a JSON path) results are looked up / stored under
    ``key`` (default: a signature of the batch's sizes and models). Returns (choice, timings)."""
    import torch
    candidates = candidates or [(b, s) for s in (1, 2) for b in (256, 512, 1024)]
    if key is None:
        ns = np.array([s["n"] for s in batch.specs])
        models = "".join(sorted({s["model"][0] for s in batch.specs}))
        key = f"{models}:{batch.nrep}:{int(ns.sum())}:{int(ns.max())}"
    if table and os.path.exists(table):
        with open(table) as f:
            cached = json.load(f)
        if key in cached:
            b, s = cached[key]["choice"]
            batch.set_launch(b, s)
            return (b, s), cached[key]["ms"]
    times = {}
    scratch = clone_state(state)
    batch.forces(scratch, energy=False)
    for b, s in candidates:
        batch.set_launch(b, s)
        work = clone_state(scratch)
        batch.run(work, 2, dt=dt, mass=mass)  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batch.run(work, steps, dt=dt, mass=mass)
        torch.cuda.synchronize()
        times[f"{b}x{s}"] = 1e3 * (time.perf_counter() - t0)
        del work
    best = min(candidates, key=lambda c: times[f"{c[0]}x{c[1]}"])
    batch.set_launch(*best)
    if table:
        cached = {}
        if os.path.exists(table):
            with open(table) as f:
                cached = json.load(f)
        cached[key] = {"choice": list(best), "ms": times, "steps": steps}
        with open(table, "w") as f:
            json.dump(cached, f, indent=1, sort_keys=True)
    return best, times
